"""Pins the dynamics half of the CPU f64 oracle: constraint clustering, SETTLE,
SHAKE, RATTLE, the LFMiddle and Verlet steps, centre-of-mass removal, kinetic
energy and degrees of freedom, the Maxwell-Boltzmann draw and the position
restraint.

Every check compares the oracle with something other than itself: a closed
form (a two-body constraint solved exactly, the leapfrog solution of a
harmonic oscillator, a geometric series, the Ornstein-Uhlenbeck variance),
an invariant (momentum, rigid distances, a rigid-body projection) or a second
method solving the same equations (SETTLE against SHAKE iterated to 1e-13,
the water velocity solve against RATTLE).  Each is chosen so that a plausible
slip -- a dropped mass factor, a sign, the current instead of the reference
bond vector, e^{-g dt} in place of e^{-2 g dt} -- fails it.  Mirrors
``test/constraints.test.ts``, ``test/integrator.test.ts``,
``test/kinetic.test.ts``, ``test/velocities.test.ts`` and
``test/restraint.test.ts`` upstream.
"""
import math

import numpy as np
import pytest

import oracle as O
from mdcore.system import (CMMotionRemover, NonbondedForce, OpenMMSystem,  # noqa: F401
                           PositionRestraintForce)

KB = 0.008314462618153242
TIP3P_OH, TIP3P_HH = 0.09572, 0.15139
M_O, M_H, M_C = 15.99943, 1.007947, 12.01078


def _system(masses, constraints=(), forces=(), box=((3.0, 0, 0), (0, 3.0, 0), (0, 0, 3.0))):
    c = np.asarray(constraints, dtype=np.float64).reshape(-1, 3)
    return OpenMMSystem(boxVectors=box, masses=np.asarray(masses, dtype=np.float64),
                        constraintP1=c[:, 0].astype(np.int32), constraintP2=c[:, 1].astype(np.int32),
                        constraintDistance=c[:, 2].copy(), forces=list(forces))


def _water(center, rng):
    """A TIP3P water at its constrained geometry, randomly oriented."""
    half = 0.5 * TIP3P_HH
    h = math.sqrt(TIP3P_OH ** 2 - half ** 2)
    local = np.array([[0.0, 0.0, 0.0], [half, h, 0.0], [-half, h, 0.0]])
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return local @ q.T + center


def _waters(count, rng, spacing=0.5):
    pos, cons, masses = [], [], []
    for w in range(count):
        c = np.array([w % 5, (w // 5) % 5, w // 25], dtype=float) * spacing + 0.3
        pos.append(_water(c, rng))
        o = 3 * w
        cons += [(o, o + 1, TIP3P_OH), (o, o + 2, TIP3P_OH), (o + 1, o + 2, TIP3P_HH)]
        masses += [M_O, M_H, M_H]
    return np.concatenate(pos).reshape(-1), cons, masses


def _methyl(rng, center):
    """A CH3 group: three C-H constraints sharing the carbon (a 4-atom cluster)."""
    d = 0.109
    dirs = rng.normal(size=(3, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    return np.vstack([center, center + d * dirs]), d


def _distances(x, cons):
    x = x.reshape(-1, 3)
    return np.array([np.linalg.norm(x[int(i)] - x[int(j)]) for i, j, _ in cons])


def _momentum(masses, v):
    return (np.asarray(masses)[:, None] * v.reshape(-1, 3)).sum(axis=0)


# ---- constraint clustering (src/engine/constraint-clusters.ts:93-249) --------

def test_split_finds_waters_by_f32_equal_sides_and_clusters_the_rest():
    rng = np.random.default_rng(1)
    wpos, wcons, wm = _waters(4, rng)
    # a "water" whose two central bonds differ below f32 resolution is still a
    # triangle; one whose bonds differ visibly is two SHAKE constraints + one
    near = [(12, 13, TIP3P_OH), (12, 14, TIP3P_OH * (1 + 1e-12)), (13, 14, TIP3P_HH)]
    far = [(15, 16, 0.1), (15, 17, 0.11), (16, 17, 0.15)]
    methyl = [(18, 19, 0.109), (18, 20, 0.109), (18, 21, 0.109)]
    masses = wm + [M_O, M_H, M_H] + [M_O, M_H, M_H] + [M_C, M_H, M_H, M_H]
    s = _system(masses, wcons + near + far + methyl)
    o = O.OracleSystem(s)
    sp = o.split()
    try:
        assert sp.n_settle == 5
        centrals = sorted(sp.settle[3 * k] for k in range(5))
        assert centrals == [0, 3, 6, 9, 12]
        for k in range(5):
            assert sp.settle_d[2 * k] == pytest.approx(TIP3P_OH, rel=1e-7)
            assert sp.settle_d[2 * k + 1] == pytest.approx(TIP3P_HH, rel=1e-7)
        iterative = sorted(sp.iterative[k] for k in range(sp.n_iterative))
        first_far = len(wcons) + len(near)
        assert iterative == list(range(first_far, first_far + 6))
        # the unequal triangle and the methyl group: 3 atoms/3 constraints and
        # 4 atoms/3 constraints
        assert (sp.max_atoms, sp.max_constraints, sp.error) == (4, 3, 0)
    finally:
        o.free_split(sp)


def test_split_flags_a_cluster_beyond_eight_atoms():
    # nine atoms in one constrained chain
    cons = [(k, k + 1, 0.1) for k in range(8)]
    o = O.OracleSystem(_system([12.0] * 9, cons))
    sp = o.split()
    try:
        assert sp.error == 1 and sp.max_atoms == 9
    finally:
        o.free_split(sp)


def test_constraints_between_frozen_atoms_are_dropped():
    o = O.OracleSystem(_system([0.0, 0.0, 12.0], [(0, 1, 0.1), (1, 2, 0.1)]))
    sp = o.split()
    try:
        assert [sp.iterative[k] for k in range(sp.n_iterative)] == [1]
    finally:
        o.free_split(sp)


# ---- SETTLE (src/reference/settle.ts:31-284) ---------------------------------

@pytest.fixture(scope="module")
def waters():
    rng = np.random.default_rng(7)
    ref, cons, masses = _waters(20, rng)
    s = _system(masses, cons)
    o = O.OracleSystem(s)
    sp = o.split()
    assert sp.n_settle == 20 and sp.n_iterative == 0
    yield o, sp, ref, cons, masses, rng
    o.free_split(sp)


def test_settle_restores_rigid_geometry_and_conserves_momentum(waters):
    o, sp, ref, cons, masses, rng = waters
    trial = ref + rng.normal(scale=0.004, size=ref.shape)
    before = trial.copy()
    o.settle_positions(sp, ref, trial)
    d = _distances(trial, cons)
    # SETTLE keeps the f32-rounded lengths the split stores, as upstream
    # (constraint-clusters.ts:169-179): 0.09572 is 5.7e-10 nm off in f32
    want = np.array([c[2] for c in cons], dtype=np.float32).astype(np.float64)
    assert np.max(np.abs(d - want) / want) < 1e-12
    # a correction along the reference bonds only: no net displacement of mass
    assert np.max(np.abs(_momentum(masses, trial - before))) < 1e-13


def test_settle_equals_shake_converged_on_the_same_equations(waters):
    o, sp, ref, cons, masses, rng = waters
    trial = ref + rng.normal(scale=0.004, size=ref.shape)
    settled = trial.copy()
    o.settle_positions(sp, ref, settled)
    # SHAKE solves the same equations (corrections along the reference bond
    # vectors, the same f32-rounded lengths): iterate it far past the usual
    # tolerance
    f32 = [(i, j, float(np.float32(d))) for i, j, d in cons]
    o2 = O.OracleSystem(_system(masses, f32))
    shaken = trial.copy()
    it, ok = o2.shake(o2.iterative_split(range(len(cons))), ref, shaken, 1e-14, max_iterations=20000)
    assert ok
    assert np.max(np.abs(shaken - settled)) < 1e-12


def test_settle_velocities_remove_bond_velocities_like_rattle(waters):
    o, sp, ref, cons, masses, rng = waters
    v = rng.normal(scale=1.0, size=ref.shape)
    vs = v.copy()
    o.settle_velocities(sp, ref, vs)
    x = ref.reshape(-1, 3)
    vv = vs.reshape(-1, 3)
    for i, j, _ in cons:
        r = x[i] - x[j]
        assert abs(np.dot(vv[i] - vv[j], r) / np.linalg.norm(r)) < 1e-12
    assert np.max(np.abs(_momentum(masses, vs - v))) < 1e-12
    vr = v.copy()
    it, ok = o.rattle(o.iterative_split(range(len(cons))), ref, vr, 1e-13, max_iterations=20000)
    assert ok
    assert np.max(np.abs(vr - vs)) < 1e-11


# ---- SHAKE and RATTLE (src/reference/constraints.ts:130-243) -----------------

def _two_body_exact(ref, trial, m1, m2, d):
    """The SHAKE solution of one constraint in closed form: trial moved along
    the reference vector r by lam (1/m1, -1/m2), |r' + lam (1/m1 + 1/m2) r| = d."""
    r = ref[0:3] - ref[3:6]
    rp = trial[0:3] - trial[3:6]
    w = 1.0 / m1 + 1.0 / m2
    a = w * w * np.dot(r, r)
    b = 2.0 * w * np.dot(rp, r)
    c = np.dot(rp, rp) - d * d
    lam = (-b + math.sqrt(b * b - 4 * a * c)) / (2 * a)  # the root near zero
    out = trial.copy()
    out[0:3] += lam * r / m1
    out[3:6] -= lam * r / m2
    return out


def test_shake_one_constraint_matches_the_closed_form():
    rng = np.random.default_rng(3)
    m1, m2, d = M_C, M_H, 0.109
    s = _system([m1, m2], [(0, 1, d)])
    o = O.OracleSystem(s)
    sp = o.split()
    try:
        for _ in range(20):
            u = rng.normal(size=3)
            ref = np.concatenate([np.zeros(3), d * u / np.linalg.norm(u)])
            trial = ref + rng.normal(scale=0.003, size=6)
            want = _two_body_exact(ref, trial, m1, m2, d)
            got = trial.copy()
            it, ok = o.shake(sp, ref, got, 1e-14, max_iterations=1000)
            assert ok and it > 0
            assert np.max(np.abs(got - want)) < 1e-14
    finally:
        o.free_split(sp)


def test_shake_methyl_conserves_momentum_and_meets_the_tolerance_window():
    rng = np.random.default_rng(4)
    x, d = _methyl(rng, np.array([1.0, 1.0, 1.0]))
    ref = x.reshape(-1)
    cons = [(0, 1, d), (0, 2, d), (0, 3, d)]
    masses = [M_C, M_H, M_H, M_H]
    o = O.OracleSystem(_system(masses, cons))
    sp = o.split()
    try:
        trial = ref + rng.normal(scale=0.003, size=ref.shape)
        before = trial.copy()
        tol = 1e-6
        it, ok = o.shake(sp, ref, trial, tol)
        assert ok
        rel = _distances(trial, cons) / d
        assert np.all(rel >= 1 - tol) and np.all(rel <= 1 + tol)
        assert np.max(np.abs(_momentum(masses, trial - before))) < 1e-14
        # each hydrogen moved only along its own reference bond (plus what
        # the carbon's motion gives every bond): the hydrogen displacements
        # minus the carbon's scaled share lie along their reference bonds
        dx = (trial - before).reshape(-1, 3)
        r = ref.reshape(-1, 3)
        for h in (1, 2, 3):
            bond = r[0] - r[h]
            perp = dx[h] - np.dot(dx[h], bond) / np.dot(bond, bond) * bond
            assert np.linalg.norm(perp) < 1e-15
    finally:
        o.free_split(sp)


def test_rattle_one_constraint_matches_the_closed_form():
    rng = np.random.default_rng(5)
    m1, m2, d = M_C, M_H, 0.109
    o = O.OracleSystem(_system([m1, m2], [(0, 1, d)]))
    sp = o.split()
    try:
        u = rng.normal(size=3)
        x = np.concatenate([np.zeros(3), d * u / np.linalg.norm(u)])
        v = rng.normal(size=6)
        got = v.copy()
        it, ok = o.rattle(sp, x, got, 1e-15, max_iterations=100)
        assert ok
        rhat = (x[0:3] - x[3:6]) / d
        vrel = np.dot(v[0:3] - v[3:6], rhat)
        mu = m1 * m2 / (m1 + m2)
        want = v.copy()
        want[0:3] -= mu / m1 * vrel * rhat
        want[3:6] += mu / m2 * vrel * rhat
        assert np.max(np.abs(got - want)) < 1e-13
    finally:
        o.free_split(sp)


# ---- integrators (src/reference/integrator.ts:68-220) ------------------------

def _oscillator(k=500.0, m=12.0):
    """Two particles, the second frozen, with a harmonic spring on the first to
    the origin expressed as a position restraint (k |x|^2 => F = -2 k x)."""
    s = _system([m], forces=[PositionRestraintForce(np.array([0], dtype=np.int32), np.array([k]),
                                                    np.zeros(1), np.zeros(1), np.zeros(1))],
                box=((50.0, 0, 0), (0, 50.0, 0), (0, 0, 50.0)))
    return s, 2.0 * k, m


@pytest.mark.parametrize("integrator", ["lfmiddle", "verlet"])
def test_step_is_the_leapfrog_solution_of_a_harmonic_oscillator(integrator):
    s, kspring, m = _oscillator()
    o = O.OracleSystem(s)
    sp = o.split()
    dt, steps = 0.002, 400
    w = math.sqrt(kspring / m)
    x = np.array([0.1, -0.05, 0.02])
    v = np.array([0.0, 0.3, -0.1])  # at t = -dt/2 (leapfrog convention)
    x0, v0 = x.copy(), v.copy()
    try:
        for n in range(steps):
            _, f = o.position_restraint(x, s.boxVectors)
            if integrator == "lfmiddle":
                o.lfmiddle_step(sp, x, v, f, dt, 0.0, 0.0, 1e-6, 1, n)
            else:
                o.verlet_step(sp, x, v, f, dt, 1e-6)
    finally:
        o.free_split(sp)
    # leapfrog on x'' = -w^2 x: x_n = A cos(n W dt) + B sin(n W dt) with
    # sin(W dt / 2) = w dt / 2 (the exact solution of the difference equation)
    W = 2.0 / dt * math.asin(w * dt / 2.0)
    x1 = x0 + dt * (v0 - dt * w * w * x0)
    A = x0
    B = (x1 - A * math.cos(W * dt)) / math.sin(W * dt)
    want = A * math.cos(W * dt * steps) + B * math.sin(W * dt * steps)
    assert np.max(np.abs(x - want)) < 1e-12


def test_lfmiddle_friction_without_noise_decays_as_a_geometric_series():
    s = _system([16.0, 1.0])
    o = O.OracleSystem(s)
    sp = o.split()
    dt, gamma, steps = 0.002, 5.0, 50
    x = np.zeros(6)
    v = np.array([1.0, -2.0, 0.5, 0.3, 0.0, -1.0])
    v0 = v.copy()
    try:
        for n in range(steps):
            o.lfmiddle_step(sp, x, v, np.zeros(6), dt, 0.0, gamma, 1e-6, 1, n)
    finally:
        o.free_split(sp)
    q = math.exp(-gamma * dt)
    assert np.max(np.abs(v - v0 * q ** steps)) < 1e-14
    # each step drifts dt/2 before and after the friction: dt/2 v (q^k + q^(k+1))
    disp = 0.5 * dt * sum(q ** k + q ** (k + 1) for k in range(steps))
    assert np.max(np.abs(x - v0 * disp)) < 1e-14


def test_lfmiddle_thermostat_has_the_ornstein_uhlenbeck_variance():
    n, m, T, gamma, dt = 20000, 4.0, 300.0, 10.0, 0.002
    s = _system([m] * n, box=((100.0, 0, 0), (0, 100.0, 0), (0, 0, 100.0)))
    o = O.OracleSystem(s)
    sp = o.split()
    x = np.zeros(3 * n)
    v = np.zeros(3 * n)
    try:
        o.lfmiddle_step(sp, x, v, np.zeros(3 * n), dt, T, gamma, 1e-6, 11, 0)
        var1 = np.mean(v * v)
        for step in range(1, 400):
            o.lfmiddle_step(sp, x, v, np.zeros(3 * n), dt, T, gamma, 1e-6, 11, step)
        var_inf = np.mean(v * v)
    finally:
        o.free_split(sp)
    kt_m = KB * T / m
    # one step from rest: (1 - e^{-2 g dt}) kT/m; after many steps: kT/m.
    # 60,000 components put the sampling error near 0.6 %
    assert var1 / ((1 - math.exp(-2 * gamma * dt)) * kt_m) == pytest.approx(1.0, abs=0.03)
    assert var_inf / kt_m == pytest.approx(1.0, abs=0.03)


def test_lfmiddle_constrained_step_keeps_constraints_and_momentum():
    rng = np.random.default_rng(9)
    ref, cons, masses = _waters(10, rng)
    x2, d = _methyl(rng, np.array([2.0, 2.0, 2.0]))
    base = len(masses)
    ref = np.concatenate([ref, x2.reshape(-1)])
    cons = cons + [(base, base + 1, d), (base, base + 2, d), (base, base + 3, d)]
    masses = masses + [M_C, M_H, M_H, M_H]
    s = _system(masses, cons)
    o = O.OracleSystem(s)
    sp = o.split()
    try:
        x = ref.copy()
        v = rng.normal(scale=0.5, size=ref.shape)
        f = rng.normal(scale=50.0, size=ref.shape)
        f -= np.repeat((f.reshape(-1, 3).sum(axis=0) / len(masses))[None, :], len(masses), axis=0).reshape(-1)
        p0 = _momentum(masses, v) + 0.002 * f.reshape(-1, 3).sum(axis=0)
        o.lfmiddle_step(sp, x, v, f, 0.002, 0.0, 0.0, 1e-10, 1, 0)
        want = np.array([c[2] for c in cons])
        want[:30] = want[:30].astype(np.float32)  # SETTLE keeps f32-rounded lengths
        assert np.max(np.abs(_distances(x, cons) - want) / want) < 1e-9
        # no friction, no noise, zero net force: constraints exchange momentum only
        assert np.max(np.abs(_momentum(masses, v) - p0)) < 1e-9
        # without friction or noise the two half drifts are one drift, so the
        # corrected velocity is exactly the constrained displacement over dt
        assert np.max(np.abs(v * 0.002 - (x - ref))) < 1e-15
    finally:
        o.free_split(sp)


# ---- COM, kinetic energy, degrees of freedom (com-motion.ts, kinetic.ts) -----

def test_com_removal_zeroes_momentum_and_leaves_frozen_atoms():
    rng = np.random.default_rng(2)
    masses = np.array([12.0, 1.0, 0.0, 16.0, 14.0])
    o = O.OracleSystem(_system(masses))
    v = rng.normal(size=15)
    v[6:9] = [0.5, 0.5, 0.5]
    before = v.copy()
    o.remove_com(v)
    assert np.max(np.abs(_momentum(masses, v))) < 1e-14
    assert np.all(v[6:9] == before[6:9])
    shift = (before - v).reshape(-1, 3)[masses > 0]
    assert np.max(np.abs(shift - shift[0])) < 1e-15  # one uniform shift
    p = _momentum(masses, before)
    assert np.allclose(shift[0], p / masses.sum(), atol=1e-15)


def test_kinetic_energy_and_degrees_of_freedom():
    masses = [12.0, 1.0, 0.0, 16.0]
    v = np.array([1.0, 0, 0, 0, 2.0, 0, 9.0, 9.0, 9.0, 0, 0, -0.5])
    o = O.OracleSystem(_system(masses, [(0, 1, 0.1)], forces=[CMMotionRemover(1)]))
    assert o.kinetic_energy(v) == pytest.approx(0.5 * (12.0 + 4.0 + 16.0 * 0.25), rel=1e-15)
    # three per moving particle, less one per constraint, less three for the
    # centre-of-mass remover (kinetic.ts:104-113)
    assert o.degrees_of_freedom() == 3 * 3 - 1 - 3
    assert O.OracleSystem(_system(masses, [(0, 1, 0.1)])).degrees_of_freedom() == 3 * 3 - 1


# ---- Maxwell-Boltzmann draw (src/reference/velocities.ts:78-117) -------------

def test_velocity_draw_is_the_domain_normals_scaled_then_projected():
    rng = np.random.default_rng(12)
    ref, cons, masses = _waters(30, rng)
    masses = list(masses)
    # 200 free particles, the first of them frozen
    s = _system(masses + [0.0] + [20.0] * 199, cons)
    n = len(masses) + 200
    pos = np.concatenate([ref, rng.uniform(0.0, 3.0, size=600)])
    o = O.OracleSystem(s)
    sp = o.split()
    try:
        T, seed = 310.0, 77
        v = o.velocities_at_temperature(sp, pos, T, seed)
        L = O.lib()
        # unconstrained particles: exactly sqrt(kT/m) times the Threefry
        # VELOCITY_DOMAIN normals, pinned by the Random123 known answers
        for i in range(len(masses) + 1, n):
            for a in range(3):
                want = math.sqrt(KB * T / 20.0) * L.orc_domain_normal(0xBB67AE85, seed, i, a)
                assert v[3 * i + a] == pytest.approx(want, rel=1e-15, abs=0)
        frozen = 3 * len(masses)
        assert np.all(v[frozen:frozen + 3] == 0.0)  # the frozen particle stays at rest
        # constrained relative velocities projected away (tolerance 1e-5)
        x = pos.reshape(-1, 3)
        vv = v.reshape(-1, 3)
        for i, j, d in cons:
            r = x[i] - x[j]
            assert abs(np.dot(vv[i] - vv[j], r) / np.linalg.norm(r)) < 1e-5
        # reproducible from its seed, different between seeds
        assert np.array_equal(v, o.velocities_at_temperature(sp, pos, T, seed))
        assert not np.array_equal(v, o.velocities_at_temperature(sp, pos, T, seed + 1))
        # the time offset subtracts (offset/m) F
        f = rng.normal(size=3 * n)
        v2 = o.velocities_at_temperature(sp, pos, T, seed, forces=f, offset=0.001)
        k = len(masses) + 5
        assert v2[3 * k] == pytest.approx(v[3 * k] - 0.001 / 20.0 * f[3 * k], rel=1e-14)
    finally:
        o.free_split(sp)


def test_velocity_draw_has_the_maxwell_boltzmann_variance():
    n, m, T = 30000, 7.0, 290.0
    o = O.OracleSystem(_system([m] * n, box=((100.0, 0, 0), (0, 100.0, 0), (0, 0, 100.0))))
    sp = o.split()
    try:
        v = o.velocities_at_temperature(sp, np.zeros(3 * n), T, 3)
    finally:
        o.free_split(sp)
    assert np.mean(v * v) / (KB * T / m) == pytest.approx(1.0, abs=0.02)
    assert abs(np.mean(v)) < 0.02 * math.sqrt(KB * T / m)


# ---- position restraint (src/reference/forces/restraint.ts:38-85) -----------

def test_restraint_energy_force_and_minimum_image():
    L = 3.0
    k = np.array([1000.0, 250.0])
    x0 = np.array([0.05, 1.5])
    y0 = np.array([1.0, 1.5])
    z0 = np.array([2.95, 1.5])
    s = _system([12.0, 12.0], forces=[PositionRestraintForce(np.array([0, 1], dtype=np.int32), k, x0, y0, z0)],
                box=((L, 0, 0), (0, L, 0), (0, 0, L)))
    o = O.OracleSystem(s)
    # particle 0 sits across the x and z faces from its reference point: the
    # nearest image is (-0.1, 0.02, +0.1) away, not ~(2.9, 0.02, -2.9)
    pos = np.array([2.95, 1.02, 0.05, 1.7, 1.5, 1.3])
    e, f = o.position_restraint(pos, s.boxVectors)
    d0 = np.array([-0.1, 0.02, 0.1])
    d1 = np.array([0.2, 0.0, -0.2])
    assert e == pytest.approx(1000.0 * d0 @ d0 + 250.0 * d1 @ d1, rel=1e-12)
    assert np.allclose(f[0:3], -2000.0 * d0, rtol=1e-10, atol=1e-10)
    assert np.allclose(f[3:6], -500.0 * d1, rtol=1e-10, atol=1e-10)
    # the force bound of restraint.test.ts: never beyond 2 k times half a box diagonal
    assert np.max(np.abs(f[0:3])) <= 2 * 1000.0 * 0.5 * math.sqrt(3) * L
    # numerical derivative
    h = 1e-6
    for a in range(6):
        p, mm = pos.copy(), pos.copy()
        p[a] += h
        mm[a] -= h
        num = -(o.position_restraint(p, s.boxVectors)[0] - o.position_restraint(mm, s.boxVectors)[0]) / (2 * h)
        assert num == pytest.approx(f[a], rel=1e-6, abs=1e-6)
    # a triclinic cell: the energy is k |nearest image|^2 of the reduced cell
    tri = ((L, 0, 0), (L / 3, L * 0.9428, 0), (-L / 3, L * 0.4714, L * 0.8165))
    e2, _ = o.position_restraint(pos, tri)
    red = O.reduce_box(tri)
    want = sum(k[t] * np.sum(O.nearest_image(red, pos[3 * t:3 * t + 3] - np.array([x0[t], y0[t], z0[t]])) ** 2)
               for t in range(2))
    assert e2 == pytest.approx(want, rel=1e-12)
