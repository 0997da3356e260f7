"""Batched replicas on the GPU.

* A batch advances several replicas of different sizes, cell shapes,
  temperatures and seeds; its trajectories are bit-identical to each replica
  run alone (every partial sum is rounded to Q32.32 per list segment or term
  before it is added, so how the work is spread over waves cannot show).
* After the batched run each replica's forces are within 2e-6 relative RMS
  of the f64 oracle and its constraints hold.
* A batch can be destroyed and its replicas go on alone; a replica
  destroyed first leaves the batch refusing to run.
"""
import numpy as np
import pytest

import edge
import oracle as O
from mdcore import native, synthetic
from mdcore.engine import HipBatch, HipEngine, RunOptions, minimise

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def replicas():
    """Four replicas (2k-6k atoms), minimised and given velocities; returns
    (replica list, per-replica (positions, velocities, step))."""
    reps = synthetic.ensemble(4, seed=99, low=2000, high=6000)
    states = []
    for r in reps:
        e = HipEngine(r.built.system, r.built.state, RunOptions(seed=r.seed, temperature=r.temperature))
        minimise(e, 100, 10.0)
        e.setVelocitiesToTemperature(r.temperature, r.seed)
        e.run(20)
        states.append((e.positions(), e.velocities(), e.steps))
        e.destroy()
    return reps, states


def _engines(reps, states):
    out = []
    for r, (x, v, step) in zip(reps, states):
        e = HipEngine(r.built.system, r.built.state, RunOptions(seed=r.seed, temperature=r.temperature))
        e.setPositions(x)
        e.setVelocities(v)
        e._native.step = step
        out.append(e)
    return out


def test_batch_is_bit_identical_to_replicas_run_alone(replicas):
    reps, states = replicas
    batched = _engines(reps, states)
    alone = _engines(reps, states)
    batch = HipBatch(batched)
    steps = 123  # two captured strides of 50 and 23 launched steps
    batch.run(steps)
    for e in alone:
        e.run(steps)
    for k, (a, b) in enumerate(zip(batched, alone)):
        print(f"replica {k}: {a.particleCount} atoms, step {a.steps}")
        assert a.steps == b.steps == states[k][2] + steps
        assert np.array_equal(a.positions(), b.positions()), k
        assert np.array_equal(a.velocities(), b.velocities()), k
    batch.destroy()
    for e in batched + alone:
        e.destroy()


def test_compacted_pair_schedule_gives_the_same_bits(replicas):
    """The compacted pair schedule (in-cutoff pairs of two entries packed into
    full waves, their force factors put back in the block layout) evaluates
    the same expressions of the same inputs and sums them in the same order as
    the direct one: trajectories with rebuilds are bit-identical, each replica
    alone and the four batched."""
    reps, states = replicas
    direct = _engines(reps, states)
    compact = _engines(reps, states)
    batched = _engines(reps, states)
    for e in direct:
        e._native.set_pair_schedule(compact=False)
    for e in compact:
        e._native.set_pair_schedule(compact=True)
    batch = HipBatch(batched)
    batch._native.set_pair_schedule(compact=True)
    steps = 123
    for e in direct + compact:
        e.run(steps)
    batch.run(steps)
    for k in range(len(reps)):
        d, c, b = direct[k], compact[k], batched[k]
        print(f"replica {k}: {d.particleCount} atoms, rebuilds {d.verletFlags()['rebuilds']}")
        assert np.array_equal(d.positions(), c.positions()), k
        assert np.array_equal(d.velocities(), c.velocities()), k
        assert np.array_equal(d.positions(), b.positions()), k
        assert np.array_equal(d.velocities(), b.velocities()), k
    batch.destroy()
    for e in direct + compact + batched:
        e.destroy()


def test_batched_forces_match_oracle_and_constraints_hold(replicas):
    reps, states = replicas
    engines = _engines(reps, states)
    batch = HipBatch(engines)
    batch.run(100)
    for r, e in zip(reps, engines):
        s = r.built.system
        x = e.positions()
        _, _, f = e._native.evaluate(True)
        o = O.OracleSystem(s)
        _, _, of = o.potential(x, s.boxVectors, grid=e.mesh)
        err, _ = edge.rel_rms_edge(f, of, s, x, s.boxVectors, o.alpha())
        xx = x.reshape(-1, 3)
        d = np.linalg.norm(xx[s.constraintP1] - xx[s.constraintP2], axis=1)
        ce = float(np.max(np.abs(d - s.constraintDistance) / s.constraintDistance))
        print(f"{s.particleCount} atoms: force rel RMS {err:.2e}, constraint error {ce:.1e}")
        assert err < 2e-6
        assert ce < 2e-6
    batch.destroy()
    for e in engines:
        e.destroy()


def test_batch_profile_and_lifetime(replicas):
    reps, states = replicas
    engines = _engines(reps, states)
    batch = HipBatch(engines)
    ph = batch.profileStep(3)
    print(ph)
    for name in ("pairs", "pair_forces", "bonded", "pme_spread", "pme_fft_convolve", "pme_gather", "integrate"):
        assert ph[name] > 0.0
    batch.destroy()
    engines[0].run(5)  # a replica goes on alone after its batch
    batch = HipBatch(engines)
    engines[1].destroy()
    with pytest.raises(native.MdError):
        batch.run(1)
    batch.destroy()
    for e in engines:
        e.destroy()


def test_autotuner_picks_a_schedule_and_the_batch_goes_on(replicas):
    """md_batch_autotune times the pair kernel's schedules on the batch and
    keeps the fastest (the starting schedule is among those timed, so the
    pick is never slower); afterwards the batch runs on and its forces still
    match the oracle."""
    reps, states = replicas
    engines = _engines(reps, states)
    batch = HipBatch(engines)
    best = batch.autotune(5)
    print("autotuned schedule", best)
    assert best["candidates"] >= 9
    assert best["segment_entries"] in (8, 16, 32, 64, 128)
    assert best["pair_waves"] > 0
    assert best["ms_per_step"] <= best["default_ms_per_step"]
    batch.run(60)
    batch.destroy()
    for e, r in zip(engines, reps):
        o = O.OracleSystem(r.built.system)
        x = e.positions()
        _, _, f = e._native.evaluate(True)
        _, _, of = o.potential(x, r.built.system.boxVectors, grid=e.mesh)
        err, _ = edge.rel_rms_edge(f, of, r.built.system, x, r.built.system.boxVectors, o.alpha())
        assert err < 2e-6
        e.destroy()
