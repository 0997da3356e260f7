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
    verr = np.max(np.abs(v_dev - v))
    print("one-step max position difference (nm):", err, "velocity (nm/ps):", verr)
    assert err < 2e-7
    # the velocity carries the constraint correction (x' - x_unconstrained)/dt,
    # so the position gate over dt bounds it
    assert verr < 2e-7 / 0.002
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
    _, og, of = o.potential(e.positions(), b.system.boxVectors, grid=tuple(int(g) for g in e.mesh))
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


def test_stale_list_evaluates_every_in_cutoff_pair(prepared):
    """The gate upstream's neighbour validation sets (validate-neighbours.ts:
    1-12): a list that misses a pair gives forces a force comparison cannot
    tell apart, and a trajectory that blows up later.  Step until the list is
    as stale as the rebuild rule allows (an atom has moved more than a
    quarter of the skin since the build, and no rebuild has happened in the
    last step), then count on the device the in-cutoff pairs the pair kernel
    evaluates from that stale list: it must equal the oracle's count at the
    same positions, pairs within f32 rounding of the cutoff excepted, and the
    count a freshly built list gives."""
    from scipy.spatial import cKDTree
    b, pos, vel, step = prepared
    e = _engine(b, pos, vel, step=step, seed=13)
    e.run(1)
    skin = 0.15
    stale = None
    for k in range(60):
        before = e._native.list_flags()
        e.run(1)
        after = e._native.list_flags()
        # between a quarter and half the skin: no rebuild is pending either
        if after["rebuilds"] == before["rebuilds"] and 0.25 * skin < after["worstDisplacement"] < 0.5 * skin:
            stale = after
            break
    assert stale is not None, "the list never became stale enough"
    pairs_stale, _ = e._native.pair_census()
    assert e._native.list_flags()["rebuilds"] == stale["rebuilds"]  # the census did not rebuild
    x = e.positions()
    fresh = _engine(b, x, e.velocities(), step=e.steps)
    pairs_fresh, _ = fresh._native.pair_census()
    o = O.OracleSystem(b.system)
    d, _ = o.direct_space(x, b.system.boxVectors, with_forces=False)
    L = np.diag(np.asarray(b.system.boxVectors))
    tree = cKDTree(np.mod(x.reshape(-1, 3), L), boxsize=L)
    rc = b.system.force("NonbondedForce").cutoff
    edge = len(tree.query_pairs(rc + 2e-6)) - len(tree.query_pairs(rc - 2e-6))
    print("worst displacement", stale["worstDisplacement"], "stale census", pairs_stale, "fresh", pairs_fresh,
          "oracle", d.pairs_geometric, "edge pairs", edge)
    assert abs(pairs_stale - d.pairs_geometric) <= edge
    assert abs(pairs_stale - pairs_fresh) <= edge
    e.destroy()
    fresh.destroy()


def test_constraints_that_cannot_converge_raise_and_leave_the_step(prepared):
    """A tolerance below f64 resolution leaves every SHAKE/RATTLE cluster
    unconverged after its passes: md_run reports ConstraintsDiverged and the
    step count stays where it was.  New positions clear the condition."""
    b, pos, vel, step = prepared
    e = _engine(b, pos, vel, step=step, constraintTolerance=1e-18)
    with pytest.raises(native.ConstraintsDiverged) as err:
        e.run(3)
    print(err.value)
    assert e.steps == step
    e.destroy()
    ok = _engine(b, pos, vel, step=step)
    ok.run(3)
    assert ok.steps == step + 3
    ok.destroy()


def test_list_overflow_during_a_run_stops_it_within_a_stride():
    """A box compressed to about six times the starting density overflows
    the cluster-pair list at the run's first rebuild: md_run raises
    ListOverflow after the first stride, not after all its steps, and the
    step count stays where it was."""
    b = synthetic.config(1)
    e = HipEngine(b.system, b.state, RunOptions(seed=3))
    scale = 0.55
    e.setBox(np.asarray(b.system.boxVectors) * scale)
    e.setPositions(b.state.positions * scale)
    with pytest.raises(native.ListOverflow) as err:
        e.run(500)
    print(err.value)
    import re
    at = int(re.search(r"at or before step (\d+)", str(err.value)).group(1))
    assert at < 50, at
    assert e.steps == 0
    e.destroy()


def test_device_minimiser_follows_the_host_search():
    """md_minimise runs the host minimiser's search (the same L-BFGS, line
    search, springs and rounds) with every vector on the device: over the
    first iterations the two paths agree closely (only summation order
    differs), and run to convergence it lowers the energy as far and leaves
    the constraints within tolerance."""
    b = synthetic.config(0)
    host = HipEngine(b.system, b.state, RunOptions())
    dev = HipEngine(b.system, b.state, RunOptions())
    start, _ = host.potential()
    h3 = minimise(host, 3, 10.0)
    d3 = dev._native.minimise(3, 10.0)
    print("3 iterations: host", h3.energy, "device", d3["energy"])
    assert abs(d3["energy"] - h3.energy) < 1e-6 * abs(h3.energy) + 1e-3
    host.destroy()
    dev.destroy()
    host = HipEngine(b.system, b.state, RunOptions())
    dev = HipEngine(b.system, b.state, RunOptions())
    hr = minimise(host, 0, 10.0)
    energy = dev.minimise(0, 10.0)
    dr = dev.lastMinimisation
    print("converged: host", hr, "device", dr, "start", start)
    assert energy < start
    assert dr.maxConstraintError < 1e-5
    assert not dr.gaveUp
    assert abs(energy - hr.energy) < 0.02 * abs(hr.energy)
    x = dev.positions().reshape(-1, 3)
    s = b.system
    r = np.linalg.norm(x[s.constraintP1] - x[s.constraintP2], axis=1)
    assert np.max(np.abs(r - s.constraintDistance) / s.constraintDistance) < 1e-5
    host.destroy()
    dev.destroy()
