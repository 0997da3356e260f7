"""Monte Carlo barostat, driven from the host around device force evaluations.

The move (src/reference/barostat.ts:210-274): propose dV ~ U(-s, s), scale
the box by (V'/V)^(1/3) and every molecule's centroid by the same factor
(atoms translate rigidly with their molecule), evaluate the energy again and
accept when

    w = dE + P dV - N_molecules kT ln(V'/V) <= 0,  or with probability e^(-w/kT).

Every ten attempts the proposal half-width s shrinks by 1.1 below 25%
acceptance and grows by 1.1 (capped at 0.3 V) above 75%.  The attempt fires
on the frequency-th step, not the zeroth (barostat.ts:129-139), and the
deviates are Threefry draws on their own domain with a draw counter carried
by the engine (src/gpu/loop.ts:602-641).  The PME mesh and the cell
divisions stay as they were built; the device recomputes every box constant.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .system import OpenMMSystem
from .units import BOLTZMANN_KJ_PER_MOL_K

# src/reference/barostat.ts:48 (the barostat's own expression for a bar)
BAR_TO_KJ_PER_MOL_NM3 = 6.02214076e23 * 1e-25


@dataclass
class Molecules:
    """Connected components of the bonded/constraint graph
    (src/engine/molecules.ts:71): ``moleculeOf[atom]``, numbered by each
    molecule's lowest atom; ``starts``/``atoms`` group the atoms ascending."""

    moleculeOf: np.ndarray
    count: int
    starts: np.ndarray
    atoms: np.ndarray
    largest: int


def find_molecules(system: OpenMMSystem) -> Molecules:
    """Edges from constraints, bonds, angles, torsions and nonbonded
    exceptions; every one of them is intramolecular, so their union gives
    the molecules whatever subset a system carries (a rigid water has only
    constraints and exceptions)."""
    n = system.particleCount
    parent = np.arange(n, dtype=np.int64)

    def find(x):
        root = x
        while parent[root] != root:
            root = parent[root]
        while parent[x] != root:  # compress the path walked
            parent[x], x = root, parent[x]
        return root

    def union_all(a, b):
        for x, y in zip(np.asarray(a, dtype=np.int64).tolist(), np.asarray(b, dtype=np.int64).tolist()):
            rx, ry = find(x), find(y)
            if rx != ry:
                if rx < ry:
                    parent[ry] = rx
                else:
                    parent[rx] = ry

    union_all(system.constraintP1, system.constraintP2)
    for f in system.forces:
        t = f.type
        if t == "HarmonicBondForce":
            union_all(f.p1, f.p2)
        elif t == "HarmonicAngleForce":
            union_all(f.p1, f.p2)
            union_all(f.p2, f.p3)
        elif t == "PeriodicTorsionForce":
            union_all(f.p1, f.p2)
            union_all(f.p2, f.p3)
            union_all(f.p3, f.p4)
        elif t == "NonbondedForce":
            union_all(f.exceptionP1, f.exceptionP2)
    roots = np.array([find(i) for i in range(n)], dtype=np.int64)
    # number molecules by their lowest atom: first appearance in atom order
    _, first, inverse = np.unique(roots, return_index=True, return_inverse=True)
    rank = np.empty(first.shape[0], dtype=np.int64)
    rank[np.argsort(first, kind="stable")] = np.arange(first.shape[0])
    molecule_of = rank[inverse].astype(np.int32)
    count = int(first.shape[0])
    sizes = np.bincount(molecule_of, minlength=count)
    starts = np.zeros(count + 1, dtype=np.int32)
    np.cumsum(sizes, out=starts[1:])
    atoms = np.argsort(molecule_of, kind="stable").astype(np.int32)
    return Molecules(molecule_of, count, starts, atoms, int(sizes.max()) if count else 0)


def scale_molecules(positions: np.ndarray, molecules: Molecules, scale: float) -> np.ndarray:
    """Each molecule's unweighted centroid times ``scale``, its atoms
    translated rigidly by the centroid's move (barostat.ts:148-177).
    Returns new (3N,) positions."""
    x = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    m = molecules.moleculeOf
    sizes = np.bincount(m, minlength=molecules.count).astype(np.float64)
    centre = np.stack([np.bincount(m, weights=x[:, k], minlength=molecules.count) for k in range(3)], axis=1)
    mean = centre / sizes[:, None]
    offset = scale * mean - mean
    return (x + offset[m]).reshape(-1)


def box_volume(box) -> float:
    """Product of the diagonal: the volume of a reduced (lower-triangular)
    box (barostat.ts:94-98)."""
    b = np.asarray(box, dtype=np.float64)
    return float(b[0, 0] * b[1, 1] * b[2, 2])


@dataclass
class BarostatState:
    """Calls since the last attempt, the proposal half-width (nm^3) and the
    adaptation window's counts (barostat.ts:52-70)."""

    step: int
    volumeScale: float
    attempted: int = 0
    accepted: int = 0

    @classmethod
    def initial(cls, box) -> "BarostatState":
        return cls(step=0, volumeScale=0.01 * box_volume(box))


def acts_after_steps(state: BarostatState, frequency: int, steps: int) -> bool:
    """Advance the call counter by ``steps``; true once it reaches the
    frequency, carrying any overshoot (barostat.ts:129-139)."""
    if frequency == 0:
        return False
    state.step += steps
    if state.step < frequency:
        return False
    state.step -= frequency
    return True


@dataclass
class AttemptResult:
    accepted: bool
    deltaVolume: float
    lengthScale: float
    initialEnergy: float
    finalEnergy: float
    work: float
    box: np.ndarray
    positions: np.ndarray


def attempt(barostat, state: BarostatState, box, positions, potential, uniform,
            molecules: Molecules) -> AttemptResult:
    """One Metropolis volume move.  ``potential(positions, box)`` returns
    the potential energy; ``uniform()`` a deviate in (0, 1).  The state's
    counters and proposal width are updated; the inputs are not."""
    box = np.asarray(box, dtype=np.float64)
    volume = box_volume(box)
    initial = potential(positions, box)
    delta = state.volumeScale * 2.0 * (uniform() - 0.5)
    new_volume = volume + delta
    length_scale = float(np.cbrt(new_volume / volume))
    trial_box = box * length_scale
    trial = scale_molecules(positions, molecules, length_scale)
    final = potential(trial, trial_box)
    pressure = barostat.defaultPressure * BAR_TO_KJ_PER_MOL_NM3
    kT = BOLTZMANN_KJ_PER_MOL_K * barostat.defaultTemperature
    work = final - initial + pressure * delta - molecules.count * kT * math.log(new_volume / volume)
    # an uphill move draws a second deviate; a downhill one does not
    accepted = not (work > 0 and uniform() > math.exp(-work / kT))
    state.attempted += 1
    if accepted:
        state.accepted += 1
    if state.attempted >= 10:
        if state.accepted < 0.25 * state.attempted:
            state.volumeScale /= 1.1
            state.attempted = state.accepted = 0
        elif state.accepted > 0.75 * state.attempted:
            state.volumeScale = min(state.volumeScale * 1.1, volume * 0.3)
            state.attempted = state.accepted = 0
    return AttemptResult(accepted, delta, length_scale, initial, final, work,
                         trial_box if accepted else box,
                         trial if accepted else np.asarray(positions, dtype=np.float64).copy())
