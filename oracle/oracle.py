"""ctypes access to the CPU f64 oracle (``liboracle.so``, built from this folder).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, by ``__graft_entry__.smoke``
and by ``bench.py``'s CPU-baseline leg, and only as the checker.  The engine
never imports it.  ``OracleSystem`` packs an ``OpenMMSystem`` (the host's
parsed System contract) into the C struct ``orc_system`` of ``oracle.h``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

G_BOND, G_ANGLE, G_TORSION, G_DIRECT, G_RECIPROCAL, G_RESTRAINT = range(6)
GROUP_NAMES = ["HarmonicBondForce", "HarmonicAngleForce", "PeriodicTorsionForce",
               "NonbondedForce[direct]", "NonbondedForce[reciprocal]", "PositionRestraint"]


def build() -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class _System(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("mass", _dp),
        ("nc", C.c_int), ("c1", _ip), ("c2", _ip), ("cd", _dp),
        ("nb", C.c_int), ("b1", _ip), ("b2", _ip), ("blen", _dp), ("bk", _dp),
        ("na", C.c_int), ("a1", _ip), ("a2", _ip), ("a3", _ip), ("atheta", _dp), ("ak", _dp),
        ("nt", C.c_int), ("t1", _ip), ("t2", _ip), ("t3", _ip), ("t4", _ip), ("tper", _ip),
        ("tphase", _dp), ("tk", _dp),
        ("has_nonbonded", C.c_int), ("cutoff", C.c_double), ("ewald_tolerance", C.c_double),
        ("alpha", C.c_double), ("grid", C.c_int * 3), ("dispersion", C.c_int),
        ("q", _dp), ("sigma", _dp), ("epsilon", _dp),
        ("ne", C.c_int), ("e1", _ip), ("e2", _ip), ("eq", _dp), ("esig", _dp), ("eeps", _dp),
        ("nr", C.c_int), ("rp", _ip), ("rk", _dp), ("rx0", _dp), ("ry0", _dp), ("rz0", _dp),
        ("cmm_frequency", C.c_int),
    ]


class _Direct(C.Structure):
    _fields_ = [("coulomb", C.c_double), ("lennard_jones", C.c_double),
                ("exclusion_correction", C.c_double), ("exceptions", C.c_double),
                ("dispersion", C.c_double), ("total", C.c_double),
                ("pairs_within_cutoff", C.c_int64), ("pairs_geometric", C.c_int64)]


class _Split(C.Structure):
    _fields_ = [("n_settle", C.c_int), ("settle", _ip), ("settle_d", _dp),
                ("n_iterative", C.c_int), ("iterative", _ip),
                ("max_atoms", C.c_int), ("max_constraints", C.c_int), ("error", C.c_int)]


def _declare(L):
    P = C.POINTER
    L.orc_to_fixed.argtypes = [C.c_float]
    L.orc_to_fixed.restype = C.c_int64
    L.orc_from_fixed_f32.argtypes = [C.c_int64]
    L.orc_from_fixed_f32.restype = C.c_float
    L.orc_from_fixed.argtypes = [C.c_int64]
    L.orc_from_fixed.restype = C.c_double
    L.orc_threefry2x32.argtypes = [P(C.c_uint32), P(C.c_uint32), C.c_int, P(C.c_uint32)]
    L.orc_open_unit_interval.argtypes = [C.c_uint32]
    L.orc_open_unit_interval.restype = C.c_double
    L.orc_thermostat_normal.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int]
    L.orc_thermostat_normal.restype = C.c_double
    L.orc_domain_normal.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int]
    L.orc_domain_normal.restype = C.c_double
    L.orc_ewald_alpha.argtypes = [C.c_double, C.c_double]
    L.orc_ewald_alpha.restype = C.c_double
    L.orc_pme_minimum_grid.argtypes = [C.c_double, C.c_double, C.c_double]
    L.orc_smooth_fft_size.argtypes = [C.c_int]
    for name in ("orc_reduce_box", "orc_inverse_box", "orc_perpendicular_widths"):
        getattr(L, name).argtypes = [_dp, _dp]
    L.orc_box_volume.argtypes = [_dp]
    L.orc_box_volume.restype = C.c_double
    L.orc_nearest_image.argtypes = [_dp, _dp]
    L.orc_is_reduced.argtypes = [_dp, C.c_double]
    L.orc_alpha_for.argtypes = [P(_System)]
    L.orc_alpha_for.restype = C.c_double
    L.orc_default_grid.argtypes = [P(_System), _dp, C.c_double, _ip]
    for name in ("orc_dispersion_correction", "orc_dispersion_correction_naive"):
        getattr(L, name).argtypes = [P(_System), C.c_double]
        getattr(L, name).restype = C.c_double
    L.orc_self_energy.argtypes = [P(_System), C.c_double]
    L.orc_self_energy.restype = C.c_double
    L.orc_net_charge_correction.argtypes = [P(_System), C.c_double, C.c_double]
    L.orc_net_charge_correction.restype = C.c_double
    L.orc_direct_space.argtypes = [P(_System), _dp, _dp, C.c_double, C.c_int, _dp, P(_Direct)]
    L.orc_bspline.argtypes = [C.c_double, C.c_int, _dp, _dp]
    L.orc_bspline_moduli.argtypes = [C.c_int, C.c_int, _dp]
    L.orc_fft3d.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.orc_dft3d.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.orc_pme_reciprocal.argtypes = [P(_System), _dp, _dp, C.c_double, _ip, _dp]
    L.orc_pme_reciprocal.restype = C.c_double
    L.orc_ewald_reciprocal.argtypes = [_dp, C.c_int, _dp, _dp, C.c_double, C.c_double, _dp,
                                       P(C.c_int64)]
    L.orc_ewald_reciprocal.restype = C.c_double
    for name in ("orc_harmonic_bond", "orc_harmonic_angle", "orc_periodic_torsion"):
        getattr(L, name).argtypes = [P(_System), _dp, _dp]
        getattr(L, name).restype = C.c_double
    L.orc_position_restraint.argtypes = [P(_System), _dp, _dp, _dp]
    L.orc_position_restraint.restype = C.c_double
    L.orc_potential.argtypes = [P(_System), _dp, _dp, _ip, _dp, _dp]
    L.orc_potential.restype = C.c_double
    L.orc_split_constraints.argtypes = [P(_System), P(_Split)]
    L.orc_split_free.argtypes = [P(_Split)]
    L.orc_settle_positions.argtypes = [P(_System), P(_Split), _dp, _dp]
    L.orc_settle_velocities.argtypes = [P(_System), P(_Split), _dp, _dp]
    L.orc_shake.argtypes = [P(_System), P(_Split), _dp, _dp, C.c_double, C.c_int, _ip]
    L.orc_rattle.argtypes = [P(_System), P(_Split), _dp, _dp, C.c_double, C.c_int, _ip]
    L.orc_remove_com.argtypes = [P(_System), _dp]
    L.orc_kinetic_energy.argtypes = [P(_System), _dp]
    L.orc_kinetic_energy.restype = C.c_double
    L.orc_degrees_of_freedom.argtypes = [P(_System)]
    L.orc_lfmiddle_step.argtypes = [P(_System), P(_Split), _dp, _dp, _dp, C.c_double, C.c_double,
                                    C.c_double, C.c_double, C.c_uint32, C.c_uint32]
    L.orc_verlet_step.argtypes = [P(_System), P(_Split), _dp, _dp, _dp, C.c_double, C.c_double]
    L.orc_velocities_at_temperature.argtypes = [P(_System), P(_Split), _dp, _dp, C.c_double,
                                                C.c_uint32, C.c_double, _dp]
    L.orc_run_lfmiddle.argtypes = [P(_System), _dp, _dp, _dp, _ip, C.c_int, C.c_double,
                                   C.c_double, C.c_double, C.c_double, C.c_uint32, C.c_uint32]
    L.orc_run_lfmiddle.restype = C.c_double


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _box9(box) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(box, dtype=np.float64).reshape(9))


class OracleSystem:
    """An ``OpenMMSystem`` packed for the C oracle; owns the arrays it points at."""

    def __init__(self, system, grid=None):
        self.system = system
        self._keep = []
        s = _System()

        def f64(x):
            a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
            self._keep.append(a)
            return _d(a)

        def i32(x):
            a = np.ascontiguousarray(np.asarray(x, dtype=np.int32))
            self._keep.append(a)
            return _i(a)

        empty_f, empty_i = np.zeros(1), np.zeros(1, dtype=np.int32)
        self._keep += [empty_f, empty_i]
        for name in ("mass", "cd", "blen", "bk", "atheta", "ak", "tphase", "tk", "q", "sigma",
                     "epsilon", "eq", "esig", "eeps", "rk", "rx0", "ry0", "rz0"):
            setattr(s, name, _d(empty_f))
        for name in ("c1", "c2", "b1", "b2", "a1", "a2", "a3", "t1", "t2", "t3", "t4", "tper",
                     "e1", "e2", "rp"):
            setattr(s, name, _i(empty_i))
        s.n = system.particleCount
        s.mass = f64(system.masses)
        s.nc = system.constraintCount
        if s.nc:
            s.c1, s.c2, s.cd = (i32(system.constraintP1), i32(system.constraintP2),
                                f64(system.constraintDistance))
        for f in system.forces:
            if f.type == "HarmonicBondForce" and f.count:
                s.nb, s.b1, s.b2, s.blen, s.bk = f.count, i32(f.p1), i32(f.p2), f64(f.length), f64(f.k)
            elif f.type == "HarmonicAngleForce" and f.count:
                s.na = f.count
                s.a1, s.a2, s.a3 = i32(f.p1), i32(f.p2), i32(f.p3)
                s.atheta, s.ak = f64(f.angle), f64(f.k)
            elif f.type == "PeriodicTorsionForce" and f.count:
                s.nt = f.count
                s.t1, s.t2, s.t3, s.t4 = i32(f.p1), i32(f.p2), i32(f.p3), i32(f.p4)
                s.tper, s.tphase, s.tk = i32(f.periodicity), f64(f.phase), f64(f.k)
            elif f.type == "NonbondedForce":
                s.has_nonbonded = 1
                s.cutoff, s.ewald_tolerance, s.alpha = f.cutoff, f.ewaldTolerance, f.alpha
                g = grid if grid is not None else f.grid
                for a in range(3):
                    s.grid[a] = int(g[a])
                s.dispersion = int(bool(f.useDispersionCorrection))
                s.q, s.sigma, s.epsilon = f64(f.charge), f64(f.sigma), f64(f.epsilon)
                s.ne = f.exceptionCount
                if s.ne:
                    s.e1, s.e2 = i32(f.exceptionP1), i32(f.exceptionP2)
                    s.eq, s.esig = f64(f.exceptionChargeProd), f64(f.exceptionSigma)
                    s.eeps = f64(f.exceptionEpsilon)
            elif f.type == "PositionRestraint" and f.count:
                s.nr, s.rp, s.rk = f.count, i32(f.particle), f64(f.k)
                s.rx0, s.ry0, s.rz0 = f64(f.x0), f64(f.y0), f64(f.z0)
            elif f.type == "CMMotionRemover":
                s.cmm_frequency = int(f.frequency)
        self.c = s
        self.ref = C.byref(s)

    # ---- forces and energies ------------------------------------------
    def alpha(self) -> float:
        return lib().orc_alpha_for(self.ref)

    def default_grid(self, box) -> tuple:
        g = np.zeros(3, dtype=np.int32)
        lib().orc_default_grid(self.ref, _d(_box9(box)), self.alpha(), _i(g))
        return tuple(int(x) for x in g)

    def potential(self, positions, box, grid=None):
        """(total, groups[6], forces[3N]) as src/reference/potential.ts:94."""
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        forces = np.zeros(3 * self.c.n)
        groups = np.zeros(6)
        g = np.zeros(3, dtype=np.int32) if grid is None else np.asarray(grid, dtype=np.int32)
        total = lib().orc_potential(self.ref, _d(pos), _d(_box9(box)), _i(g), _d(forces), _d(groups))
        return total, groups, forces

    def direct_space(self, positions, box, alpha=None, naive=False, with_forces=True):
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        forces = np.zeros(3 * self.c.n)
        out = _Direct()
        a = self.alpha() if alpha is None else alpha
        lib().orc_direct_space(self.ref, _d(pos), _d(_box9(box)), a, int(naive),
                               _d(forces) if with_forces else None, C.byref(out))
        return out, forces

    def pme_reciprocal(self, positions, box, grid, alpha=None):
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        forces = np.zeros(3 * self.c.n)
        g = np.asarray(grid, dtype=np.int32)
        a = self.alpha() if alpha is None else alpha
        e = lib().orc_pme_reciprocal(self.ref, _d(pos), _d(_box9(box)), a, _i(g), _d(forces))
        return e, forces

    def self_energy(self, alpha=None):
        return lib().orc_self_energy(self.ref, self.alpha() if alpha is None else alpha)

    def net_charge_correction(self, box, alpha=None):
        return lib().orc_net_charge_correction(self.ref, self.alpha() if alpha is None else alpha,
                                               lib().orc_box_volume(_d(_box9(box))))

    def dispersion(self, box, naive=False):
        v = lib().orc_box_volume(_d(_box9(box)))
        f = lib().orc_dispersion_correction_naive if naive else lib().orc_dispersion_correction
        return f(self.ref, v)

    def bonded(self, positions):
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        out = {}
        for key, fn in (("bond", lib().orc_harmonic_bond), ("angle", lib().orc_harmonic_angle),
                        ("torsion", lib().orc_periodic_torsion)):
            f = np.zeros(3 * self.c.n)
            out[key] = (fn(self.ref, _d(pos), _d(f)), f)
        return out

    # ---- constraints and dynamics --------------------------------------
    def split(self):
        sp = _Split()
        lib().orc_split_constraints(self.ref, C.byref(sp))
        return sp

    def free_split(self, sp):
        lib().orc_split_free(C.byref(sp))

    def iterative_split(self, constraints):
        """A split that sends the given constraint indices to SHAKE/RATTLE and
        none to SETTLE (not freed with ``free_split``: numpy owns it)."""
        idx = np.ascontiguousarray(constraints, dtype=np.int32)
        self._keep.append(idx)
        sp = _Split()
        sp.n_iterative = int(idx.shape[0])
        sp.iterative = _i(idx)
        return sp

    def settle_positions(self, sp, reference, trial):
        """SETTLE every rigid water of ``trial`` (in place) against ``reference``."""
        lib().orc_settle_positions(self.ref, C.byref(sp), _d(reference), _d(trial))

    def settle_velocities(self, sp, positions, velocities):
        lib().orc_settle_velocities(self.ref, C.byref(sp), _d(positions), _d(velocities))

    def shake(self, sp, reference, trial, tolerance, max_iterations=150):
        """SHAKE the iterative clusters of ``trial`` in place: (iterations, converged)."""
        conv = C.c_int(0)
        it = lib().orc_shake(self.ref, C.byref(sp), _d(reference), _d(trial), tolerance,
                             max_iterations, C.byref(conv))
        return int(it), bool(conv.value)

    def rattle(self, sp, positions, velocities, tolerance, max_iterations=150):
        conv = C.c_int(0)
        it = lib().orc_rattle(self.ref, C.byref(sp), _d(positions), _d(velocities), tolerance,
                              max_iterations, C.byref(conv))
        return int(it), bool(conv.value)

    def lfmiddle_step(self, sp, positions, velocities, forces, dt, temperature, friction,
                      tolerance, seed, step):
        lib().orc_lfmiddle_step(self.ref, C.byref(sp), _d(positions), _d(velocities), _d(forces),
                                dt, temperature, friction, tolerance, seed, step)

    def verlet_step(self, sp, positions, velocities, forces, dt, tolerance):
        lib().orc_verlet_step(self.ref, C.byref(sp), _d(positions), _d(velocities), _d(forces), dt,
                              tolerance)

    def velocities_at_temperature(self, sp, positions, temperature, seed, forces=None, offset=0.0):
        v = np.zeros(3 * self.c.n)
        f = None if forces is None else _d(np.ascontiguousarray(forces, dtype=np.float64))
        lib().orc_velocities_at_temperature(self.ref, C.byref(sp),
                                            _d(np.ascontiguousarray(positions, dtype=np.float64)), f,
                                            temperature, seed, offset, _d(v))
        return v

    def position_restraint(self, positions, box):
        """(energy, forces) of the PositionRestraint terms alone."""
        f = np.zeros(3 * self.c.n)
        e = lib().orc_position_restraint(self.ref, _d(np.ascontiguousarray(positions, dtype=np.float64)),
                                         _d(_box9(box)), _d(f))
        return e, f

    def remove_com(self, velocities):
        lib().orc_remove_com(self.ref, _d(velocities))

    def kinetic_energy(self, velocities):
        return lib().orc_kinetic_energy(self.ref, _d(np.ascontiguousarray(velocities)))

    def degrees_of_freedom(self):
        return lib().orc_degrees_of_freedom(self.ref)

    def run_lfmiddle(self, positions, velocities, box, steps, dt, temperature, friction,
                     tolerance, seed, first_step=0, grid=None):
        g = np.zeros(3, dtype=np.int32) if grid is None else np.asarray(grid, dtype=np.int32)
        return lib().orc_run_lfmiddle(self.ref, _d(positions), _d(velocities), _d(_box9(box)),
                                      _i(g), steps, dt, temperature, friction, tolerance, seed,
                                      first_step)


def threefry(counter, key, rounds=20):
    c = (C.c_uint32 * 2)(*counter)
    k = (C.c_uint32 * 2)(*key)
    out = (C.c_uint32 * 2)()
    lib().orc_threefry2x32(c, k, rounds, out)
    return int(out[0]), int(out[1])


def to_fixed(value: float) -> int:
    return int(lib().orc_to_fixed(value))


def from_fixed_f32(fixed: int) -> float:
    return float(lib().orc_from_fixed_f32(fixed))


def from_fixed(fixed: int) -> float:
    return float(lib().orc_from_fixed(fixed))


def fft3d(data: np.ndarray, sign: int, direct: bool = False) -> np.ndarray:
    """In-place unnormalised complex transform of a (nx,ny,nz) complex128 array."""
    a = np.ascontiguousarray(data, dtype=np.complex128).copy()
    nx, ny, nz = a.shape
    flat = a.view(np.float64).reshape(-1)
    fn = lib().orc_dft3d if direct else lib().orc_fft3d
    fn(_d(flat), nx, ny, nz, sign)
    return a


def bspline(w: float, order: int = 5):
    wt, dw = np.zeros(order), np.zeros(order)
    lib().orc_bspline(w, order, _d(wt), _d(dw))
    return wt, dw


def bspline_moduli(n: int, order: int = 5) -> np.ndarray:
    m = np.zeros(n)
    lib().orc_bspline_moduli(n, order, _d(m))
    return m


def ewald_reciprocal(charges, positions, box, alpha, tolerance=1e-10):
    q = np.ascontiguousarray(charges, dtype=np.float64)
    pos = np.ascontiguousarray(positions, dtype=np.float64)
    forces = np.zeros_like(pos)
    vec = C.c_int64()
    e = lib().orc_ewald_reciprocal(_d(q), q.shape[0], _d(pos), _d(_box9(box)), alpha, tolerance,
                                   _d(forces), C.byref(vec))
    return e, forces, vec.value


def nearest_image(box, d):
    v = np.ascontiguousarray(d, dtype=np.float64).copy()
    lib().orc_nearest_image(_d(_box9(box)), _d(v))
    return v


def reduce_box(box):
    out = np.zeros(9)
    lib().orc_reduce_box(_d(_box9(box)), _d(out))
    return out.reshape(3, 3)
