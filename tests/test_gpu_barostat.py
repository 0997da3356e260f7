"""The Monte Carlo barostat on the GPU engine against the oracle.

* after setBox with molecule-scaled positions (the mesh fixed), device forces
  are within 2e-6 relative RMS and group energies within 1e-5 of the oracle
  at the same box and mesh (src/gpu/loop.ts:1031-1097);
* one Metropolis attempt with the device's potential and one with the
  oracle's, on the same deviates, agree on the decision and on the work;
* an NPT run attempts every 25 steps, keeps every constraint and leaves the
  device box equal to the host's.
"""
import numpy as np
import pytest

import edge
import oracle as O
from mdcore import barostat as baro
from mdcore import synthetic
from mdcore.engine import HipEngine, RunOptions, minimise

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def npt():
    b = synthetic.build(300, 700, 2, "cube", 77, barostat=True)
    e = HipEngine(b.system, b.state, RunOptions(seed=4))
    minimise(e, 200, 10.0)
    e.setVelocitiesToTemperature(300.0, 4)
    e.run(100)
    yield b, e
    e.destroy()


def _rel_rms(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def test_forces_after_set_box_match_oracle(npt):
    b, e = npt
    x = e.positions()
    box = np.asarray(e.boxVectors())
    mol = baro.find_molecules(b.system)
    y = baro.scale_molecules(x, mol, 1.02)
    e.setBox(box * 1.02)
    e.setPositions(y)
    total, groups, f = e._native.evaluate(True)
    o = O.OracleSystem(b.system)
    ot, og, of = o.potential(y, box * 1.02, grid=e.mesh)
    err, _ = edge.rel_rms_edge(f, of, b.system, y, box * 1.02, o.alpha())
    gerr = [abs(groups[k] - og[k]) / max(abs(og[k]), 1.0) for k in range(6)]
    print("set_box: mesh", e.mesh, "force rel rms", err, "group errors", gerr)
    e.setBox(box)
    e.setPositions(x)
    assert err < 2e-6
    assert max(gerr) < 1e-5


def test_attempt_agrees_with_oracle_attempt(npt):
    b, e = npt
    x = e.positions()
    box = np.asarray(e.boxVectors())
    mol = baro.find_molecules(b.system)
    o = O.OracleSystem(b.system)
    mesh = e.mesh

    def gpu_potential(trial, trial_box):
        e.setBox(trial_box)
        e.setPositions(trial)
        return e._native.evaluate(False)[0]

    def cpu_potential(trial, trial_box):
        return o.potential(np.asarray(trial), trial_box, grid=mesh)[0]

    for draws in ([0.3, 0.5], [0.8, 0.2]):
        sg = baro.BarostatState.initial(box)
        sc = baro.BarostatState.initial(box)
        ig, ic = iter(draws), iter(draws)
        rg = baro.attempt(b.system.force("MonteCarloBarostat"), sg, box, x, gpu_potential, lambda: next(ig), mol)
        rc = baro.attempt(b.system.force("MonteCarloBarostat"), sc, box, x, cpu_potential, lambda: next(ic), mol)
        print("attempt: gpu work", rg.work, "oracle work", rc.work, "accepted", rg.accepted, rc.accepted)
        assert rg.accepted == rc.accepted
        assert abs(rg.work - rc.work) < 1e-5 * abs(rc.initialEnergy)
        assert abs(rg.initialEnergy - rc.initialEnergy) < 1e-5 * abs(rc.initialEnergy)
    e.setBox(box)
    e.setPositions(x)


def test_npt_run_attempts_on_schedule_and_keeps_constraints(npt):
    b, e = npt
    start_box = np.asarray(e.boxVectors())
    before = e.barostatAttempts
    e.run(250)
    assert e.barostatAttempts - before == 10
    x = e.positions().reshape(-1, 3)
    s = b.system
    r = np.linalg.norm(x[s.constraintP1] - x[s.constraintP2], axis=1)
    rel = np.abs(r - s.constraintDistance) / s.constraintDistance
    dev_box = e._native.box()
    print("attempts", e.barostatAttempts, "accepted", e.barostatAccepted, "box", np.diag(dev_box),
          "from", np.diag(start_box), "constraint error", rel.max())
    np.testing.assert_allclose(dev_box, e._box, rtol=0, atol=1e-12)
    assert rel.max() < 2e-6
    assert np.all(np.isfinite(x))
    assert e.barostatAccepted > 0


def _twin(b, e, **kw):
    """A second engine at e's state (positions, velocities, step, box)."""
    from mdcore.state import OpenMMState
    st = OpenMMState(positions=e.positions(), velocities=e.velocities(), boxVectors=e.boxVectors(),
                     stepCount=e.steps)
    return HipEngine(b.system, st, RunOptions(seed=4, **kw))


def test_device_attempt_matches_the_host_attempt(npt):
    """One attempt on the device (inside the captured stride) and one on the
    host (upstream's path), from the same state: the same deviates, the same
    decision, the same box and positions.  The device's cube root, logarithm
    and exponential are its own library's, so the box and the positions may
    differ by the last bit; everything before them is the same arithmetic."""
    b, e = npt
    dev = _twin(b, e, barostat="device")
    host = _twin(b, e, barostat="host")
    for k in range(3):
        dev.run(25)
        host.run(25)
        d, h = dev.lastBarostat, host.lastBarostat
        xd, xh = dev.positions(), host.positions()
        bd, bh = np.asarray(dev.boxVectors()), np.asarray(host.boxVectors())
        print("device", d["last_accepted"], d["last_energy_before"], d["last_energy_after"], d["last_work"],
              "host", h["last_accepted"], h["last_energy_before"], h["last_energy_after"], h["last_work"],
              "box diff", np.max(np.abs(bd - bh)), "position diff", np.max(np.abs(xd - xh)))
        assert d["last_accepted"] == h["last_accepted"]
        assert d["draws"] == h["draws"]
        assert abs(d["last_work"] - h["last_work"]) < 1e-6 * abs(h["last_energy_before"])
        if k == 0:
            # the same state and arithmetic up to the library functions
            assert abs(d["last_energy_before"] - h["last_energy_before"]) < 1e-12 * abs(h["last_energy_before"])
            np.testing.assert_allclose(bd, bh, rtol=1e-14, atol=0)
            assert np.max(np.abs(xd - xh)) < 1e-8
        else:
            # a last-bit difference in a box length moves f32 forces by their
            # last bit, which 25 steps of dynamics grow but do not amplify far
            assert abs(d["last_energy_before"] - h["last_energy_before"]) < 1e-6 * abs(h["last_energy_before"])
            np.testing.assert_allclose(bd, bh, rtol=1e-10, atol=0)
            assert np.max(np.abs(xd - xh)) < 1e-5
    assert dev.barostatAttempts == host.barostatAttempts == 3
    dev.destroy()
    host.destroy()


def test_device_barostat_graphs_equal_direct_launches(npt):
    """The attempt inside captured strides and launched kernel by kernel (no
    graphs) gives the same bits, over runs whose lengths cut the schedule
    at different places."""
    b, e = npt
    g = _twin(b, e)
    d = _twin(b, e, maxBatch=1)
    for n in (30, 45, 25):
        g.run(n)
        d.run(n)
    assert g.barostatAttempts == d.barostatAttempts == 4
    assert np.array_equal(g.positions(), d.positions())
    assert np.array_equal(np.asarray(g.boxVectors()), np.asarray(d.boxVectors()))
    assert g.barostatState.step == d.barostatState.step == 0
    g.destroy()
    d.destroy()


def test_device_barostat_leaves_tables_matching_its_box(npt):
    """After device attempts have changed the box, the forces at the final
    state match the oracle's at the box the engine reports (every subsystem's
    box constants were rewritten on the device), and the constraints hold."""
    b, e = npt
    t = _twin(b, e)
    start = np.asarray(t.boxVectors())
    t.run(200)
    box = np.asarray(t.boxVectors())
    x = t.positions()
    _, groups, f = t._native.evaluate(True)
    o = O.OracleSystem(b.system)
    _, og, of = o.potential(x, box, grid=t.mesh)
    err, _ = edge.rel_rms_edge(f, of, b.system, x, box, o.alpha())
    gerr = [abs(groups[k] - og[k]) / max(abs(og[k]), 1.0) for k in range(6)]
    s = b.system
    xr = x.reshape(-1, 3)
    r = np.linalg.norm(xr[s.constraintP1] - xr[s.constraintP2], axis=1)
    rel = np.abs(r - s.constraintDistance) / s.constraintDistance
    print("device barostat: attempts", t.barostatAttempts, "accepted", t.barostatAccepted, "volume",
          np.prod(np.diag(start)), "->", np.prod(np.diag(box)), "force rel rms", err, "groups", gerr,
          "constraints", rel.max())
    assert t.barostatAttempts == 8
    assert t.barostatAccepted > 0
    assert not np.array_equal(box, start)
    assert err < 2e-6
    assert max(gerr) < 1e-5
    assert rel.max() < 2e-6
    t.destroy()
