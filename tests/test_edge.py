"""The edge-aware force comparison (tests/edge.py) on a planted pair.

Two charged atoms straddle the cutoff by 1e-8 nm: the helper must find the
pair, say which side the f64 oracle put it on, and -- given "device" forces
that counted it the other way -- recover a zero error; a pair 1e-5 nm from
the cutoff is not an edge pair, and a genuine force error is not absorbed.
"""
import math

import numpy as np

import edge
from mdcore.system import NonbondedForce, OpenMMSystem


def _two_atoms(r):
    nb = NonbondedForce(charge=np.array([-0.8, 0.4]), sigma=np.array([0.3, 0.25]), epsilon=np.array([0.6, 0.1]),
                        exceptionP1=np.zeros(0, np.int32), exceptionP2=np.zeros(0, np.int32),
                        exceptionChargeProd=np.zeros(0), exceptionSigma=np.zeros(0), exceptionEpsilon=np.zeros(0),
                        cutoff=1.0)
    box = ((4.0, 0.0, 0.0), (0.0, 4.0, 0.0), (0.0, 0.0, 4.0))
    system = OpenMMSystem(boxVectors=box, masses=np.array([16.0, 1.0]), constraintP1=np.zeros(0, np.int32),
                          constraintP2=np.zeros(0, np.int32), constraintDistance=np.zeros(0), forces=[nb])
    x = np.array([1.0, 1.0, 1.0, 1.0 + r, 1.0, 1.0])
    return system, x, box


ALPHA = math.sqrt(-math.log(2 * 5e-4))


def test_planted_edge_pair_is_found_and_absorbed():
    system, x, box = _two_atoms(1.0 + 1e-8)
    pairs = edge.edge_pairs(system, x, box, ALPHA)
    assert len(pairs) == 1
    i, j, fj, inside = pairs[0]
    assert (i, j) == (0, 1) and not inside
    # opposite charges attract: the force on j points back towards i (-x)
    assert fj[0] < 0
    oracle = np.zeros((2, 3)) + 5.0  # any background force
    device = oracle.copy()
    device[1] += fj  # the device counted the pair
    device[0] -= fj
    err, n = edge.rel_rms_edge(device, oracle, system, x, box, ALPHA)
    assert n == 1 and err < 1e-15
    # a real discrepancy elsewhere survives
    device[0, 1] += 0.5
    err, _ = edge.rel_rms_edge(device, oracle, system, x, box, ALPHA)
    assert err > 0.99 * 0.5 / np.linalg.norm(oracle)


def test_pairs_off_the_edge_are_left_alone():
    system, x, box = _two_atoms(1.0 + 1e-5)
    assert edge.edge_pairs(system, x, box, ALPHA) == []
