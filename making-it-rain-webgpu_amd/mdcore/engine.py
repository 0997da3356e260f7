"""``HipEngine``: the upstream ``ProtocolEngine`` over the gfx950 C ABI.

Mirrors ``LoopEngine`` (src/gpu/engine.ts:16-104) over ``StepLoop``
(src/gpu/loop.ts:208-1116): same method names and meanings, so the protocol
(src/engine/protocol.ts) drives it unchanged.  The device owns the state
between calls; the host computes what upstream computes on the host in f64:
the reported kinetic energy and temperature (src/reference/kinetic.ts), the
Maxwell-Boltzmann draw (src/reference/velocities.ts:78-117) and the L-BFGS
minimiser with constraint springs (src/reference/minimise.ts:164-280).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import barostat as baro
from . import native
from .state import OpenMMState
from .system import OpenMMSystem, reduceBox
from .units import BOLTZMANN_KJ_PER_MOL_K

THERMOSTAT_DOMAIN = 0x9E3779B9
BAROSTAT_DOMAIN = 0x6A09E667
VELOCITY_DOMAIN = 0xBB67AE85
VELOCITY_CONSTRAINT_TOLERANCE = 1e-5   # src/reference/velocities.ts:56
SHIFTED_VELOCITY_TOLERANCE = 1e-4      # src/reference/kinetic.ts:28
DEFAULT_VERLET_SKIN = 0.15             # src/gpu/nonbonded.ts:227


@dataclass
class RunOptions:
    """``RunOptions`` (src/gpu/loop.ts:117-168)."""

    timeStep: float = 0.002
    temperature: float = 300.0
    friction: float = 1.0
    constraintTolerance: float = 1e-6
    seed: int = 0
    integrator: str = "LangevinMiddle"  # or "Verlet"
    verletSkin: float = DEFAULT_VERLET_SKIN
    mesh: Optional[Sequence[int]] = None
    device: int = 0
    maxBatch: int = 0  # steps per captured HIP graph; 0 = 50, 1 = no graphs
    # where a MonteCarloBarostat's moves run: "device" (inside the captured
    # strides, md_barostat_attach) or "host" (between strides, as upstream)
    barostat: str = "device"


@dataclass
class Report:
    """``Report`` (src/gpu/loop.ts:173-199), with per-force-group energies."""

    step: int
    time: float
    potentialEnergy: float
    kineticEnergy: float
    totalEnergy: float
    temperature: float
    groups: dict = field(default_factory=dict)


# ---- counter-based RNG on the host (src/reference/random.ts) --------------

def threefry2x32(c0, c1, k0, k1, rounds: int = 20):
    """Vectorised Threefry-2x32 over uint32 arrays."""
    c0 = np.asarray(c0, dtype=np.uint32)
    c1 = np.asarray(c1, dtype=np.uint32)
    k0 = np.uint32(k0)
    k1 = np.uint32(k1)
    ks = (k0, k1, np.uint32(0x1BD11BDA) ^ k0 ^ k1)
    rot = (13, 15, 26, 6, 17, 29, 16, 24)
    with np.errstate(over="ignore"):
        x0 = c0 + ks[0]
        x1 = c1 + ks[1]
        for r in range(rounds):
            x0 = x0 + x1
            b = rot[r % 8]
            x1 = ((x1 << np.uint32(b)) | (x1 >> np.uint32(32 - b))) ^ x0
            if (r + 1) % 4 == 0:
                inj = (r + 1) // 4
                x0 = x0 + ks[inj % 3]
                x1 = x1 + ks[(inj + 1) % 3] + np.uint32(inj)
    return x0, x1


def _unit(bits):
    return ((bits >> np.uint32(9)).astype(np.float64) + 0.5) * 2.0 ** -23


def domain_normals(domain: int, seed: int, n: int) -> np.ndarray:
    """domainNormal(domain, seed, particle, axis) for every particle: (n, 3)."""
    p = np.arange(n, dtype=np.uint64)
    out = np.zeros((n, 3))
    for half in (0, 1):
        c0 = ((2 * p + half) & 0xFFFFFFFF).astype(np.uint32)
        b0, b1 = threefry2x32(c0, np.full(n, seed & 0xFFFFFFFF, np.uint32), seed & 0xFFFFFFFF, domain)
        r = np.sqrt(-2.0 * np.log(_unit(b0)))
        a = 2.0 * math.pi * _unit(b1)
        if half == 0:
            out[:, 0], out[:, 1] = r * np.cos(a), r * np.sin(a)
        else:
            out[:, 2] = r * np.cos(a)
    return out


# ---- the engine ------------------------------------------------------------

class HipEngine:
    """``ProtocolEngine`` (src/engine/protocol.ts:32-62) on one MI355X."""

    def __init__(self, system: OpenMMSystem, state: OpenMMState, options: Optional[RunOptions] = None):
        self.system = system
        self.options = options or RunOptions()
        o = self.options
        self._box = reduceBox(state.boxVectors)
        self._native = native.NativeSystem(
            system, time_step=o.timeStep, temperature=o.temperature, friction=o.friction,
            constraint_tolerance=o.constraintTolerance, seed=o.seed, integrator=o.integrator,
            verlet_skin=o.verletSkin, mesh=o.mesh or (0, 0, 0), device=o.device, box=self._box,
            max_batch=o.maxBatch)
        self._native.set_positions(state.positions)
        v = state.velocities if state.velocities is not None else np.zeros_like(state.positions)
        self._native.set_velocities(v)
        self._native.step = int(state.stepCount)
        m = np.asarray(system.masses)
        self._mass = m
        self._moving = m > 0
        cmm = system.force("CMMotionRemover")
        self.degreesOfFreedom = int(3 * np.count_nonzero(self._moving) - system.constraintCount
                                    - (3 if cmm is not None else 0))
        # Monte Carlo barostat (src/gpu/loop.ts:602-685), deviates from their
        # own Threefry domain: on the device inside the strides, or host-driven
        # moves between strides
        self.barostat = system.force("MonteCarloBarostat")
        self.barostatState = baro.BarostatState.initial(self._box) if self.barostat is not None else None
        self.barostatDraw = 0
        self.barostatAttempts = 0
        self.barostatAccepted = 0
        self._molecules = None
        self.lastBarostat = None
        if o.barostat not in ("device", "host"):
            raise ValueError(f"barostat must be 'device' or 'host', not {o.barostat!r}")
        self._device_barostat = (self.barostat is not None and self.barostat.frequency > 0
                                 and o.barostat == "device")
        if self._device_barostat:
            self._molecules = baro.find_molecules(system)
            b = self.barostat
            self._native.attach_barostat(b.defaultPressure, b.defaultTemperature, b.frequency,
                                         self._molecules.starts, self._molecules.atoms, seed=o.seed,
                                         volume_scale=self.barostatState.volumeScale)

    # -- ProtocolEngine ----------------------------------------------------
    @property
    def particleCount(self) -> int:
        return self.system.particleCount

    @property
    def steps(self) -> int:
        return self._native.step

    def positions(self) -> np.ndarray:
        return self._native.positions()

    def velocities(self) -> np.ndarray:
        return self._native.velocities()

    def boxVectors(self):
        return tuple(tuple(float(x) for x in r) for r in self._native.box())

    @property
    def mesh(self):
        """The PME mesh, fixed for the engine's life (src/gpu/loop.ts:973-981)."""
        return self._native.mesh()

    def boxLengths(self):
        b = self._native.box()
        return (float(b[0, 0]), float(b[1, 1]), float(b[2, 2]))

    def setPositions(self, positions) -> None:
        self._native.set_positions(positions)

    def setVelocities(self, velocities) -> None:
        self._native.set_velocities(velocities)

    def setBox(self, box) -> None:
        b = np.asarray(box, dtype=np.float64)
        if b.shape == (3,):
            b = np.diag(b)
        self._box = reduceBox(b)
        self._native.set_box(self._box)

    def run(self, steps: int) -> None:
        """``steps`` steps; with a barostat, cut at each attempt so the move
        happens on the step OpenMM puts it on (src/gpu/loop.ts:662-685)."""
        steps = int(steps)
        b = self.barostat
        if b is None or b.frequency <= 0:
            self._native.run(steps)
            return
        if self._device_barostat:
            try:
                self._native.run(steps)
            finally:
                self._read_device_barostat()
            return
        done = 0
        while done < steps:
            until = b.frequency - self.barostatState.step
            chunk = min(max(1, until), steps - done)
            self._native.run(chunk)
            done += chunk
            self._attempt_barostat(chunk)

    def _read_device_barostat(self) -> None:
        """The device barostat's counters and box into the host mirror."""
        st = self._native.barostat_state()
        self._box = reduceBox(self._native.box())
        self.barostatDraw = int(st["draws"])
        self.barostatAttempts = int(st["attempts"])
        self.barostatAccepted = int(st["accepted"])
        s = self.barostatState
        s.step = int(st["steps_since_attempt"])
        s.volumeScale = float(st["volume_scale"])
        s.attempted = int(st["window_attempted"])
        s.accepted = int(st["window_accepted"])
        self.lastBarostat = st

    def _attempt_barostat(self, steps_just_run: int) -> None:
        if not baro.acts_after_steps(self.barostatState, self.barostat.frequency, steps_just_run):
            return
        if self._molecules is None:
            self._molecules = baro.find_molecules(self.system)
        seed = int(self.options.seed) & 0xFFFFFFFF

        def uniform():
            b0, _ = threefry2x32([self.barostatDraw & 0xFFFFFFFF], [seed], BAROSTAT_DOMAIN, 0)
            self.barostatDraw += 1
            return float(_unit(b0)[0])

        def potential(trial, trial_box):
            self.setBox(trial_box)
            self._native.set_positions(trial)
            total, _, _ = self._native.evaluate(False)
            return total

        res = baro.attempt(self.barostat, self.barostatState, self._box, self.positions(), potential, uniform,
                           self._molecules)
        # the trial is on the device either way: commit the winner
        self.setBox(res.box)
        self.setPositions(res.positions)
        self.barostatAttempts += 1
        if res.accepted:
            self.barostatAccepted += 1
        self.lastBarostat = {"last_accepted": int(res.accepted), "last_energy_before": res.initialEnergy,
                             "last_energy_after": res.finalEnergy, "last_work": res.work,
                             "last_length_scale": res.lengthScale, "draws": self.barostatDraw}

    def sample(self) -> dict:
        r = self.report()
        return {"potentialEnergy": r.potentialEnergy, "kineticEnergy": r.kineticEnergy,
                "totalEnergy": r.totalEnergy, "temperature": r.temperature}

    def minimise(self, maxIterations: int = 0, tolerance: float = 10.0) -> float:
        """ProtocolEngine.minimise: the device-resident minimiser (md_minimise);
        the module's ``minimise`` runs the same search on the host."""
        r = self._native.minimise(maxIterations, tolerance)
        self.lastMinimisation = MinimiseResult(r["energy"], r["iterations"], r["rounds"], bool(r["gave_up"]),
                                               r["max_constraint_error"])
        return r["energy"]

    # -- StepLoop extras ------------------------------------------------------
    def potential(self, positions=None):
        """(energy, forces) at `positions` (or the current ones)."""
        if positions is not None:
            self._native.set_positions(positions)
        total, _, forces = self._native.evaluate(True)
        return total, forces

    def evaluateForces(self):
        return self._native.evaluate(True)

    def kineticEnergy(self, velocities=None) -> float:
        """1/2 sum m v^2, unshifted (the LangevinMiddle convention)."""
        v = (self.velocities() if velocities is None else velocities).reshape(-1, 3)
        return 0.5 * float(np.sum(self._mass[:, None] * v * v))

    def shiftedKineticEnergy(self, forces=None) -> float:
        """The Verlet convention: v + (dt/2) F/m, re-constrained at 1e-4
        (src/reference/kinetic.ts:59-84)."""
        v = self.velocities().reshape(-1, 3)
        if forces is None:
            _, _, forces = self._native.evaluate(True)
        inv = np.where(self._moving, 1.0 / np.where(self._moving, self._mass, 1.0), 0.0)
        shifted = v + 0.5 * self.options.timeStep * forces.reshape(-1, 3) * inv[:, None]
        shifted = rattle_host(self.system, self.positions().reshape(-1, 3), shifted, inv,
                              SHIFTED_VELOCITY_TOLERANCE)
        return 0.5 * float(np.sum(self._mass[:, None] * shifted * shifted))

    def report(self) -> Report:
        total, groups, forces = self._native.evaluate(True)
        if self.options.integrator == "Verlet":
            ke = self.shiftedKineticEnergy(forces)
        else:
            ke = self.kineticEnergy()
        temp = 2.0 * ke / (self.degreesOfFreedom * BOLTZMANN_KJ_PER_MOL_K) if self.degreesOfFreedom > 0 else 0.0
        step = self.steps
        return Report(step=step, time=step * self.options.timeStep, potentialEnergy=total, kineticEnergy=ke,
                      totalEnergy=total + ke, temperature=temp,
                      groups={name: float(g) for name, g in zip(native.MD_GROUP_NAMES, groups)})

    def verletFlags(self) -> dict:
        return self._native.list_flags()

    def advanceStepCount(self, steps: int) -> None:
        self._native.step = self._native.step + int(steps)

    def profileStep(self, repeats: int = 10) -> dict:
        return self._native.profile_steps(repeats)

    def setVelocitiesToTemperature(self, temperature: float, seed: int) -> None:
        """Maxwell-Boltzmann draw with the -(dt/2) F/m offset of a leapfrog
        integrator, projected at 1e-5 (src/reference/velocities.ts:78-117)."""
        n = self.particleCount
        z = domain_normals(VELOCITY_DOMAIN, seed, n)
        kT = BOLTZMANN_KJ_PER_MOL_K * temperature
        m = np.where(self._moving, self._mass, 1.0)
        v = np.sqrt(kT / m)[:, None] * z
        _, _, forces = self._native.evaluate(True)
        v -= (0.5 * self.options.timeStep / m)[:, None] * forces.reshape(-1, 3)
        v[~self._moving] = 0.0
        self._native.set_velocities(v.reshape(-1))
        self._native.constrain_velocities(VELOCITY_CONSTRAINT_TOLERANCE)

    def destroy(self) -> int:
        self._native.close()
        return 0


class HipBatch:
    """Several ``HipEngine`` replicas advanced together: every phase of a
    step is one kernel launch over all of them and strides are captured as
    HIP graphs.  Each replica keeps its own options (temperature, seed, ...)
    and all of its own methods between batch runs; the batch borrows the
    engines, which must outlive it.  No counterpart upstream (one system per
    device there); replicas under a Monte Carlo barostat are refused, since
    its moves are host-driven per system."""

    def __init__(self, engines: Sequence[HipEngine]):
        self.engines = list(engines)
        for e in self.engines:
            if e.barostat is not None and e.barostat.frequency > 0:
                raise ValueError("a batched replica cannot carry a Monte Carlo barostat")
        self._native = native.NativeBatch([e._native for e in self.engines])

    def run(self, steps: int) -> None:
        self._native.run(int(steps))

    def profileStep(self, repeats: int = 10) -> dict:
        return self._native.profile_steps(repeats)

    def autotune(self, steps: int = 10) -> dict:
        """Pick the pair kernel's schedule for this batch by timing (no effect
        on the trajectories' bits); see ``md_batch_autotune``."""
        return self._native.autotune(steps)

    @property
    def particleCount(self) -> int:
        return sum(e.particleCount for e in self.engines)

    def destroy(self) -> None:
        self._native.close()


def rattle_host(system: OpenMMSystem, pos: np.ndarray, vel: np.ndarray, inv: np.ndarray, tol: float,
                max_iterations: int = 150) -> np.ndarray:
    """Jacobi RATTLE over every constraint (src/reference/constraints.ts:195-243)."""
    p1 = np.asarray(system.constraintP1)
    p2 = np.asarray(system.constraintP2)
    if p1.size == 0:
        return vel
    v = vel.copy()
    rv = pos[p1] - pos[p2]
    l2 = np.sum(rv * rv, axis=1)
    s = inv[p1] + inv[p2]
    mu = np.where(s == 0, 0.0, 0.5 / np.where(s == 0, 1.0, s))
    for _ in range(max_iterations):
        proj = np.sum((v[p1] - v[p2]) * rv, axis=1)
        delta = np.where(l2 == 0, 0.0, -2.0 * mu * proj / np.where(l2 == 0, 1.0, l2))
        if np.all(np.abs(delta) <= tol):
            break
        dr = rv * delta[:, None]
        np.add.at(v, p1, dr * inv[p1, None])
        np.add.at(v, p2, -dr * inv[p2, None])
    return v


@dataclass
class MinimiseResult:
    energy: float
    iterations: int
    rounds: int
    gaveUp: bool
    maxConstraintError: float


def _max_constraint_error(system, x):
    p1, p2, d = system.constraintP1, system.constraintP2, system.constraintDistance
    if len(p1) == 0:
        return 0.0
    r = np.linalg.norm(x[p1] - x[p2], axis=1)
    return float(np.max(np.abs(r - d) / d))


def lbfgs(objective, x, max_iterations: int = 0, rms_tolerance: float = 10.0, particles: int = 1,
          memory: int = 6, max_move: float = 0.5, max_line_search: int = 40, armijo: float = 1e-4,
          curvature: float = 0.9):
    """L-BFGS with a backtracking strong-Wolfe line search, the search OpenMM
    runs through libLBFGS and the reference follows (src/reference/lbfgs.ts
    :102-275): `memory` correction pairs, a first step of 1/|d| and later
    steps of 1, the step halved when sufficient decrease fails or the slope
    overshoots and multiplied by 2.1 while the slope stays steep.  Stops when
    the RMS gradient per particle is at most `rms_tolerance` (libLBFGS's
    |g|/max(|x|,1) test with OpenMM's epsilon), after `max_iterations` (0: no
    limit), or when the line search cannot make progress, keeping the last
    accepted point.  One addition: a trial point that moves any coordinate
    more than `max_move` nm from the last accepted point counts as a failed
    decrease and is not evaluated, and a step with sufficient decrease that
    could only grow past that limit is taken as it is, so the search cannot
    throw atoms across the cell from a clashed start.

    `objective(x) -> (value, gradient)`; returns (x, value, iterations)."""
    x = np.array(x, dtype=np.float64, copy=True)
    value, g = objective(x)

    def done(grad):
        return math.sqrt(float(grad @ grad) / particles) <= rms_tolerance

    if done(g):
        return x, value, 0
    S, Y, RHO = [], [], []
    d = -g
    step = 1.0 / float(np.linalg.norm(d))
    iteration = 0
    while True:
        iteration += 1
        slope0 = float(g @ d)
        if slope0 >= 0.0:
            return x, value, iteration - 1
        accepted = False
        for _ in range(max_line_search):
            trial = x + step * d
            if float(np.max(np.abs(step * d))) > max_move:
                step *= 0.5
                continue
            v, gt = objective(trial)
            if not v <= value + armijo * step * slope0:
                step *= 0.5
                continue
            slope = float(gt @ d)
            if slope < curvature * slope0:
                if 2.1 * step * float(np.max(np.abs(d))) > max_move:
                    accepted = True  # still steep, but at the displacement cap
                    break
                step *= 2.1
            elif slope > -curvature * slope0:
                step *= 0.5
            else:
                accepted = True
                break
        if not accepted:
            return x, value, iteration - 1
        s, y = trial - x, gt - g
        x, value, g = trial, v, gt
        if done(g) or (max_iterations and iteration >= max_iterations):
            return x, value, iteration
        sy = float(s @ y)
        if sy > 0.0:
            S.append(s)
            Y.append(y)
            RHO.append(1.0 / sy)
            if len(S) > memory:
                S.pop(0), Y.pop(0), RHO.pop(0)
        # two-loop recursion: d = -H g
        q = -g.copy()
        alphas = []
        for s_k, y_k, r_k in zip(reversed(S), reversed(Y), reversed(RHO)):
            a = r_k * float(s_k @ q)
            alphas.append(a)
            q -= a * y_k
        if S:
            q *= 1.0 / (RHO[-1] * float(Y[-1] @ Y[-1]))
        for (s_k, y_k, r_k), a in zip(zip(S, Y, RHO), reversed(alphas)):
            b = r_k * float(y_k @ q)
            q += (a - b) * s_k
        d = q
        step = 1.0


def minimise(engine: HipEngine, maxIterations: int = 0, tolerance: float = 10.0) -> MinimiseResult:
    """OpenMM's LocalEnergyMinimizer objective: U(x) + sum (k/2)(r - d)^2 over
    the constraints, k = 100 / max(1e-4, tol), stiffened tenfold per round
    until every constraint is within max(1e-4, tol); gives up to the starting
    positions when stiffening stops helping; then projects to the caller's
    tolerance (src/reference/minimise.ts:164-280).  The search is `lbfgs`
    above (memory 6, libLBFGS's default, which OpenMM keeps)."""
    system = engine.system
    n = system.particleCount
    ctol = engine.options.constraintTolerance
    work = max(1e-4, ctol)
    k = 100.0 / work
    engine._native.constrain_positions(work)
    x0 = engine.positions().reshape(-1, 3)
    initial = x0.copy()
    frozen = ~engine._moving
    p1, p2, dist = system.constraintP1, system.constraintP2, system.constraintDistance

    def objective(flat):
        x = flat.reshape(-1, 3)
        engine._native.set_positions(flat)
        energy, _, forces = engine._native.evaluate(True)
        g = -forces.reshape(-1, 3)
        g[frozen] = 0.0
        spring = 0.0
        if len(p1):
            d = x[p1] - x[p2]
            r = np.linalg.norm(d, axis=1)
            dr = r - dist
            spring = 0.5 * k * float(np.sum(dr * dr))
            gd = (k * dr / np.where(r == 0, 1.0, r))[:, None] * d
            np.add.at(g, p1, gd)
            np.add.at(g, p2, -gd)
        return energy + spring, g.reshape(-1)

    prev1 = prev2 = 1e10
    rounds = iterations = 0
    gave_up = False
    x = x0.reshape(-1).copy()
    while True:
        rounds += 1
        x, _, nit = lbfgs(objective, x, maxIterations, tolerance, n)
        iterations += nit
        err = _max_constraint_error(system, x.reshape(-1, 3))
        if err <= work or len(p1) == 0:
            break
        if err >= prev2:
            x = initial.reshape(-1).copy()
            gave_up = True
            break
        prev2, prev1 = prev1, err
        k *= 10.0
        if err > 100 * work:
            x = initial.reshape(-1).copy()
    engine._native.set_positions(x)
    if ctol < work:
        engine._native.constrain_positions(ctol)
    energy, _, _ = engine._native.evaluate(False)
    return MinimiseResult(energy, iterations, rounds, gave_up,
                          _max_constraint_error(system, engine.positions().reshape(-1, 3)))
