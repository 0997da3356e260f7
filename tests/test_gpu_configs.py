"""The larger benchmark configurations on the GPU, at their stated meshes,
against the f64 oracle.

* ~100k atoms in a rhombic dodecahedron, PME 96^3, heavy-atom position
  restraints at k = 1000 kJ/mol/nm^2 and a Monte Carlo barostat every 25
  steps: forces and all six force groups (restraint included, on a frame
  displaced from the restraint references) against the oracle, then 50 steps
  through the engine with the barostat attempting on its schedule.
* ~250k atoms in a 13.7 nm cube, PME 128^3: forces and groups against the
  oracle at one frame.
* Six replicas of the batched ensemble spanning its size range (2,262 to
  29,884 atoms, both cell shapes, meshes 32^3-56^3): each replica's forces
  against the oracle after a batched run.

Gates as elsewhere: forces within 2e-6 relative RMS (pairs within f32
rounding of the cutoff counted either way), groups within 1e-5 relative.
"""
import concurrent.futures as cf
import multiprocessing as mp

import numpy as np
import pytest

import edge
import oracle as O
from mdcore import native, synthetic
from mdcore.engine import HipBatch, HipEngine, RunOptions, minimise

pytestmark = pytest.mark.gpu

FORCE_RMS = 2e-6
ENERGY_REL = 1e-5


def _check(system, x, s, label):
    o = O.OracleSystem(system)
    grid = tuple(int(g) for g in system.force("NonbondedForce").grid)
    assert s.mesh() == grid
    total, groups, forces = s.evaluate()
    otot, og, of = o.potential(x, system.boxVectors, grid=grid)
    err, _ = edge.rel_rms_edge(forces, of, system, x, system.boxVectors, o.alpha())
    print(f"{label}: {system.particleCount} atoms, mesh {grid}, force rel RMS {err:.2e}, groups {groups} vs {og}")
    assert err < FORCE_RMS
    for k in range(6):
        scale = max(abs(og[k]), 1e-6 * abs(otot), 1.0)
        assert abs(groups[k] - og[k]) / scale < ENERGY_REL, (label, k, groups[k], og[k])
    return groups, og


def test_config2_dodecahedron_restraints_and_barostat():
    b = synthetic.config(2)
    system = b.system
    assert 99_000 <= system.particleCount <= 101_000
    assert system.force("PositionRestraint") is not None and system.force("MonteCarloBarostat") is not None
    # a frame displaced from the restraint references, so the restraint group is not zero
    rng = np.random.default_rng(2)
    x = b.state.positions + rng.normal(0.0, 0.004, b.state.positions.shape)
    s = native.NativeSystem(system)
    s.set_positions(x)
    groups, og = _check(system, x, s, "config 2")
    assert groups[5] > 1.0 and og[5] > 1.0
    s.close()
    e = HipEngine(system, b.state, RunOptions(timeStep=0.0005, friction=10.0, seed=17))
    minimise(e, 40, 10.0)  # the generated geometry has contacts a raw start would integrate badly
    e.run(50)  # attempts after steps 25 and 50
    assert e.barostatAttempts == 2
    x2 = e.positions()
    assert np.all(np.isfinite(x2))
    r = np.linalg.norm(x2.reshape(-1, 3)[system.constraintP1] - x2.reshape(-1, 3)[system.constraintP2], axis=1)
    assert np.max(np.abs(r - system.constraintDistance) / system.constraintDistance) < 2e-6
    e.destroy()


def test_config4_250k_atoms_at_the_128_mesh():
    b = synthetic.config(4)
    system = b.system
    assert 249_000 <= system.particleCount <= 251_000
    s = native.NativeSystem(system)
    s.set_positions(b.state.positions)
    _check(system, b.state.positions, s, "config 4")
    s.close()


def test_ensemble_replicas_across_the_size_range_after_a_batched_run():
    plan = synthetic.ensemble_plan(64)
    sizes = [p["n_chain"] + 3 * p["n_waters"] + p["n_ions"] for p in plan]
    order = np.argsort(sizes)
    pick = [int(order[k]) for k in np.linspace(0, len(order) - 1, 6).round().astype(int)]
    # spawned workers: a forked child would inherit this process's GPU context
    with cf.ProcessPoolExecutor(max_workers=6, mp_context=mp.get_context("spawn")) as pool:
        built = list(pool.map(synthetic.build_planned, [plan[k] for k in pick]))
    shapes = {plan[k]["shape"] for k in pick}
    assert built[-1].system.particleCount == max(sizes) and built[0].system.particleCount == min(sizes)
    print("replicas", [(b.system.particleCount, plan[k]["shape"], plan[k]["grid"][0]) for b, k in zip(built, pick)])
    assert shapes == {"cube", "octahedron"}
    engines = [HipEngine(b.system, b.state, RunOptions(timeStep=0.0005, friction=10.0, seed=plan[k]["seed"],
                                                      temperature=plan[k]["temperature"]))
               for b, k in zip(built, pick)]
    batch = HipBatch(engines)
    batch.run(30)
    batch.destroy()
    for e, b in zip(engines, built):
        x = e.positions()
        _check(b.system, x, e._native, "replica")
        e.destroy()
