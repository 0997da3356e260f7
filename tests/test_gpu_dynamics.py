"""Dynamics on the GPU against the oracle and the upstream gates.

* one LFMiddle step from the device's own forces within 2e-7 nm of the f64
  oracle's step (src/gpu/validate-loop.ts:193-225);
* NVE (Verlet) energy drift under 1.5 kJ/mol/atom/ns from a least-squares fit
  of the shifted-KE total energy over >= 10 reports (validate-loop.ts:90-94,
  261-276);
* SETTLE waters rigid and SHAKE clusters within tolerance after a run;
* the minimiser lowers the energy and leaves the constraints satisfied;
* a list built with a skin gives the forces a fresh build gives.
"""
import numpy as np
import pytest

import oracle as O
from mdcore import native, synthetic
from mdcore.engine import HipEngine, RunOptions, minimise

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prepared():
    """Config 0, minimised and thermalised once; (system, positions, velocities)."""
    b = synthetic.config(0)
    e = HipEngine(b.system, b.state, RunOptions(seed=11))
    minimise(e, 200, 10.0)
    e.setVelocitiesToTemperature(300.0, 11)
    e.run(200)
    out = (b, e.positions(), e.velocities(), e.steps)
    e.destroy()
    return out


def _engine(b, pos, vel, step=0, **kw):
    e = HipEngine(b.system, b.state, RunOptions(**kw))
    e.setPositions(pos)
    e.setVelocities(vel)
    e._native.step = step
    return e


def test_one_step_matches_oracle_given_device_forces(prepared):
    b, pos, vel, step = prepared
    e = _engine(b, pos, vel, step=step, seed=5)
    _, _, forces = e._native.evaluate(True)
    x_dev_start = e.positions()
    v_dev_start = e.velocities()
    e.run(1)
    x_dev, v_dev = e.positions(), e.velocities()
    o = O.OracleSystem(b.system)
    sp = o.split()
    x = x_dev_start.copy()
    v = v_dev_start.copy()
    o.remove_com(v)  # CMMotionRemover(1) acts at the start of every step
    o.lfmiddle_step(sp, x, v, forces, 0.002, 300.0, 1.0, 1e-6, 5, step)
    o.free_split(sp)
    err = np.max(np.abs(x_dev - x))
    print("one-step max position difference (nm):", err, "velocity:", np.max(np.abs(v_dev - v)))
    assert err < 2e-7
    e.destroy()


def test_constraints_hold_after_a_run(prepared):
    b, pos, vel, step = prepared
    e = _engine(b, pos, vel, step=step, seed=3)
    e.run(300)
    x = e.positions().reshape(-1, 3)
    s = b.system
    r = np.linalg.norm(x[s.constraintP1] - x[s.constraintP2], axis=1)
    rel = np.abs(r - s.constraintDistance) / s.constraintDistance
    print("worst relative constraint error:", rel.max())
    assert rel.max() < 2e-6
    e.destroy()


def test_nve_energy_drift_gate(prepared):
    b, pos, vel, step = prepared
    # equilibrate further under the thermostat first: the drift gate is a
    # property of an equilibrated system (validate-loop.ts:261-276)
    eq = _engine(b, pos, vel, step=step, seed=21)
    eq.run(3000)
    e = _engine(b, eq.positions(), eq.velocities(), step=eq.steps, integrator="Verlet", friction=0.0)
    eq.destroy()
    times, energies = [], []
    for k in range(12):
        e.run(400)
        r = e.report()
        times.append(r.time)
        energies.append(r.totalEnergy)
    t = np.asarray(times)  # ps
    en = np.asarray(energies)
    slope = np.polyfit(t, en, 1)[0]  # kJ/mol/ps
    drift = abs(slope) * 1000.0 / b.system.particleCount  # kJ/mol/atom/ns
    print("NVE drift (kJ/mol/atom/ns):", drift, "energies:", en)
    assert np.all(np.isfinite(en))
    assert drift < 1.5
    e.destroy()


def test_minimiser_lowers_energy_and_keeps_constraints():
    b = synthetic.config(0)
    e = HipEngine(b.system, b.state, RunOptions())
    before, _ = e.potential()
    res = minimise(e, 100, 10.0)
    x = e.positions().reshape(-1, 3)
    s = b.system
    r = np.linalg.norm(x[s.constraintP1] - x[s.constraintP2], axis=1)
    print("minimised", before, "->", res.energy, "constraint error", res.maxConstraintError)
    assert res.energy < before
    assert np.max(np.abs(r - s.constraintDistance) / s.constraintDistance) < 1e-5
    e.destroy()


def test_skin_list_gives_fresh_list_forces(prepared):
    b, pos, vel, step = prepared
    e = _engine(b, pos, vel, step=step, seed=9)
    e.run(30)  # the list was built at the first step with a 0.15 nm skin
    _, g1, f1 = e._native.evaluate(True)
    fresh = _engine(b, e.positions(), e.velocities(), step=e.steps)
    _, g2, f2 = fresh._native.evaluate(True)
    # the two lists group the same pairs into different f32 partial sums, so
    # they agree to rounding; a missing pair would show at the 1e-3 level
    rms = np.sqrt(np.sum((f1 - f2) ** 2) / np.sum(f2 ** 2))
    o = O.OracleSystem(b.system)
    _, og, of = o.potential(e.positions(), b.system.boxVectors, grid=o.default_grid(b.system.boxVectors))
    e1 = np.sqrt(np.sum((f1 - of) ** 2) / np.sum(of ** 2))
    e2 = np.sqrt(np.sum((f2 - of) ** 2) / np.sum(of ** 2))
    print("stale-vs-fresh list force rel RMS:", rms, "vs oracle:", e1, e2)
    assert rms < 2e-6
    assert e1 < 2e-6 and e2 < 2e-6
    assert abs(g1[3] - g2[3]) < 1e-6 * abs(g2[3])
    e.destroy()
    fresh.destroy()


def test_profile_reports_every_phase(prepared):
    b, pos, vel, step = prepared
    e = _engine(b, pos, vel, step=step)
    ph = e.profileStep(5)
    print(ph)
    for name in ("pairs", "pair_forces", "bonded", "pme_spread", "pme_fft_convolve", "pme_gather", "integrate"):
        assert ph[name] > 0.0
    assert e.steps == step + 5
    e.destroy()


def test_fault_injection_surfaces_as_an_error(prepared):
    b, pos, vel, step = prepared
    e = _engine(b, pos, vel, step=step)
    e._native.inject_fault(10)
    e.run(5)
    with pytest.raises(native.FaultInjected):
        e.run(10)
    e._native.inject_fault(-1)
    e.run(2)
    e.destroy()
