"""ctypes binding of ``libmdcore.so`` (the C ABI in ``include/mdcore.h``).

This is the reference-side binding a host would add in place of the WebGPU
device: it marshals an ``OpenMMSystem`` into ``md_system_desc``, owns the
opaque ``md_system*`` handle, and maps error codes onto the named errors the
upstream host throws (``UnsupportedForce``, ``CutoffTooLargeForBox``,
``ClusterTooLarge``, ``RangeError`` for fixed-point overflow...).  There is no
fallback: if the library is missing or no GPU is visible, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmdcore.so")

MD_GROUP_NAMES = ("HarmonicBondForce", "HarmonicAngleForce", "PeriodicTorsionForce",
                  "NonbondedForce[direct]", "NonbondedForce[reciprocal]", "PositionRestraint")

# md_err codes
MD_OK = 0
ERRORS = {
    1: "UnsupportedForce", 2: "CutoffTooLargeForBox", 3: "ClusterTooLarge",
    4: "FixedPointOverflow", 5: "HipError", 6: "OutOfMemory", 7: "InvalidArgument",
    8: "ListOverflow", 9: "NotSmooth", 10: "ConstraintsDiverged", 11: "FaultInjected",
    12: "NoDevice",
}


class MdError(RuntimeError):
    """A non-zero ``md_err``; ``name`` is the upstream error class it maps to."""

    def __init__(self, code: int, message: str):
        self.code = code
        self.name = ERRORS.get(code, f"md_err {code}")
        super().__init__(f"{self.name}: {message}")


class CutoffTooLargeForBox(MdError):
    pass


class ClusterTooLarge(MdError):
    pass


class FixedPointOverflow(MdError, ArithmeticError):
    pass


class FaultInjected(MdError):
    pass


class ConstraintsDiverged(MdError):
    pass


class ListOverflow(MdError):
    pass


class NoDevice(MdError):
    pass


_SPECIFIC = {2: CutoffTooLargeForBox, 3: ClusterTooLarge, 4: FixedPointOverflow, 8: ListOverflow,
             10: ConstraintsDiverged, 11: FaultInjected, 12: NoDevice}

_ip = C.POINTER(C.c_int32)
_dp = C.POINTER(C.c_double)


class SystemDesc(C.Structure):
    _fields_ = [
        ("n_particles", C.c_int32), ("masses", _dp), ("box", C.c_double * 9),
        ("n_constraints", C.c_int32), ("constraint_p1", _ip), ("constraint_p2", _ip),
        ("constraint_distance", _dp),
        ("n_bonds", C.c_int32), ("bond_p1", _ip), ("bond_p2", _ip), ("bond_length", _dp), ("bond_k", _dp),
        ("n_angles", C.c_int32), ("angle_p1", _ip), ("angle_p2", _ip), ("angle_p3", _ip),
        ("angle_theta", _dp), ("angle_k", _dp),
        ("n_torsions", C.c_int32), ("torsion_p1", _ip), ("torsion_p2", _ip), ("torsion_p3", _ip),
        ("torsion_p4", _ip), ("torsion_periodicity", _ip), ("torsion_phase", _dp), ("torsion_k", _dp),
        ("has_nonbonded", C.c_int32), ("cutoff", C.c_double), ("ewald_tolerance", C.c_double),
        ("alpha", C.c_double), ("grid", C.c_int32 * 3), ("use_dispersion_correction", C.c_int32),
        ("charge", _dp), ("sigma", _dp), ("epsilon", _dp),
        ("n_exceptions", C.c_int32), ("exception_p1", _ip), ("exception_p2", _ip),
        ("exception_chargeprod", _dp), ("exception_sigma", _dp), ("exception_epsilon", _dp),
        ("n_restraints", C.c_int32), ("restraint_particle", _ip), ("restraint_k", _dp),
        ("restraint_x0", _dp), ("restraint_y0", _dp), ("restraint_z0", _dp),
        ("cmm_frequency", C.c_int32),
        ("has_barostat", C.c_int32), ("barostat_pressure", C.c_double),
        ("barostat_temperature", C.c_double), ("barostat_frequency", C.c_int32),
        ("barostat_seed", C.c_uint32),
    ]


class RunOptionsC(C.Structure):
    _fields_ = [("time_step", C.c_double), ("temperature", C.c_double), ("friction", C.c_double),
                ("constraint_tolerance", C.c_double), ("seed", C.c_uint32), ("integrator", C.c_int32),
                ("verlet_skin", C.c_double), ("mesh", C.c_int32 * 3), ("device", C.c_int32),
                ("max_batch", C.c_int32)]


class Schedule(C.Structure):
    _fields_ = [("segment_entries", C.c_int32), ("pair_waves", C.c_int32), ("ms_per_step", C.c_double),
                ("default_segment_entries", C.c_int32), ("default_pair_waves", C.c_int32),
                ("default_ms_per_step", C.c_double), ("candidates", C.c_int32), ("compact_pairs", C.c_int32),
                ("default_compact_pairs", C.c_int32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class MinimiseResultC(C.Structure):
    _fields_ = [("energy", C.c_double), ("iterations", C.c_int32), ("rounds", C.c_int32), ("gave_up", C.c_int32),
                ("max_constraint_error", C.c_double)]


class BarostatDescC(C.Structure):
    _fields_ = [("pressure", C.c_double), ("temperature", C.c_double), ("frequency", C.c_int32),
                ("n_molecules", C.c_int32), ("molecule_start", C.POINTER(C.c_int32)),
                ("molecule_atoms", C.POINTER(C.c_int32)), ("seed", C.c_uint32), ("volume_scale", C.c_double),
                ("draws", C.c_int64), ("steps_since_attempt", C.c_int32), ("window_attempted", C.c_int32),
                ("window_accepted", C.c_int32)]


class BarostatStateC(C.Structure):
    _fields_ = [("attempts", C.c_int64), ("accepted", C.c_int64), ("draws", C.c_int64),
                ("steps_since_attempt", C.c_int32), ("window_attempted", C.c_int32),
                ("window_accepted", C.c_int32), ("volume_scale", C.c_double), ("last_accepted", C.c_int32),
                ("last_energy_before", C.c_double), ("last_energy_after", C.c_double), ("last_work", C.c_double),
                ("last_length_scale", C.c_double)]


class ListFlags(C.Structure):
    _fields_ = [("rebuilds", C.c_int64), ("tiles", C.c_int64), ("tile_capacity", C.c_int64),
                ("worst_displacement", C.c_double), ("overflowed", C.c_int32)]


_lib = None


def library():
    """The loaded ``libmdcore.so``; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() or "
                              "`make -C making-it-rain-webgpu_amd`")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.md_last_error.restype = C.c_char_p
        L.md_version.restype = C.c_char_p
        L.md_system_create.argtypes = [P(SystemDesc), P(RunOptionsC), P(C.c_void_p)]
        L.md_system_destroy.argtypes = [C.c_void_p]
        L.md_system_destroy.restype = None
        for name in ("md_set_positions", "md_set_velocities", "md_get_positions", "md_get_velocities",
                     "md_set_box", "md_get_box"):
            getattr(L, name).argtypes = [C.c_void_p, _dp]
        L.md_evaluate.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.md_run.argtypes = [C.c_void_p, C.c_int64]
        L.md_get_step.argtypes = [C.c_void_p, P(C.c_int64)]
        L.md_set_step.argtypes = [C.c_void_p, C.c_int64]
        L.md_list_flags_get.argtypes = [C.c_void_p, P(ListFlags)]
        L.md_pair_census.argtypes = [C.c_void_p, P(C.c_int64), P(C.c_int64)]
        L.md_probe_fixed_point.argtypes = [P(C.c_float), C.c_int32, P(C.c_int64), P(C.c_float)]
        L.md_probe_thermostat.argtypes = [C.c_uint32, C.c_uint32, C.c_int32, P(C.c_uint32), P(C.c_float)]
        L.md_inject_fault.argtypes = [C.c_void_p, C.c_int32]
        L.md_profile_steps.argtypes = [C.c_void_p, C.c_int32, _dp]
        L.md_phase_name.argtypes = [C.c_int32]
        L.md_phase_name.restype = C.c_char_p
        L.md_constrain_positions.argtypes = [C.c_void_p, C.c_double]
        L.md_constrain_velocities.argtypes = [C.c_void_p, C.c_double]
        L.md_get_mesh.argtypes = [C.c_void_p, _ip]
        L.md_probe_cell_order.argtypes = [_ip, _ip]
        L.md_probe_ewald_fit.argtypes = [C.c_double, C.c_double, P(C.c_float), P(C.c_float), P(C.c_float),
                                         P(C.c_int32), _dp, _dp]
        L.md_batch_create.argtypes = [P(C.c_void_p), C.c_int32, P(C.c_void_p)]
        L.md_batch_destroy.argtypes = [C.c_void_p]
        L.md_batch_destroy.restype = None
        L.md_batch_run.argtypes = [C.c_void_p, C.c_int64]
        L.md_batch_profile_steps.argtypes = [C.c_void_p, C.c_int32, _dp]
        L.md_autotune.argtypes = [C.c_void_p, C.c_int32, C.POINTER(Schedule)]
        L.md_minimise.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.POINTER(MinimiseResultC)]
        L.md_batch_autotune.argtypes = [C.c_void_p, C.c_int32, C.POINTER(Schedule)]
        L.md_set_pair_schedule.argtypes = [C.c_void_p, C.POINTER(Schedule)]
        L.md_batch_set_pair_schedule.argtypes = [C.c_void_p, C.POINTER(Schedule)]
        L.md_barostat_attach.argtypes = [C.c_void_p, C.POINTER(BarostatDescC)]
        L.md_barostat_get.argtypes = [C.c_void_p, C.POINTER(BarostatStateC)]
        _fp = P(C.c_float)
        L.md_warmstart_create.argtypes = [C.c_int32, C.c_uint32, C.c_int64, P(C.c_void_p)]
        L.md_warmstart_destroy.argtypes = [C.c_void_p]
        L.md_warmstart_destroy.restype = None
        L.md_warmstart_param_count.restype = C.c_int32
        L.md_warmstart_record.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
        L.md_warmstart_sample_count.argtypes = [C.c_void_p, P(C.c_int64)]
        L.md_warmstart_train.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_double, _dp]
        L.md_warmstart_attach.argtypes = [C.c_void_p, C.c_void_p]
        L.md_warmstart_params.argtypes = [C.c_void_p, _fp, _fp]
        L.md_warmstart_evaluate.argtypes = [C.c_void_p, _fp, _fp, C.c_int32, _fp, _fp, _dp]
        L.md_shake_counts.argtypes = [C.c_void_p, P(C.c_int64), P(C.c_int64)]
        _lib = L
    return _lib


def check(code: int) -> None:
    if code != MD_OK:
        msg = library().md_last_error().decode(errors="replace")
        raise _SPECIFIC.get(code, MdError)(code, msg)


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    return a.ctypes.data_as(_ip)


def describe_fields(system, box: Optional[Sequence] = None) -> dict:
    """``md_system_desc``'s fields for an OpenMMSystem, by name: numbers, and
    numpy arrays (float64 / int32) for the arrays.  The same record a Node
    host hands the N-API binding (``node/engine.js``)."""
    f64 = lambda x: np.ascontiguousarray(np.asarray(x, dtype=np.float64))  # noqa: E731
    i32 = lambda x: np.ascontiguousarray(np.asarray(x, dtype=np.int32))  # noqa: E731
    d = {"n_particles": system.particleCount, "masses": f64(system.masses),
         "box": f64(np.asarray(box if box is not None else system.boxVectors, dtype=np.float64).reshape(9)),
         "n_constraints": system.constraintCount}
    if system.constraintCount:
        d.update(constraint_p1=i32(system.constraintP1), constraint_p2=i32(system.constraintP2),
                 constraint_distance=f64(system.constraintDistance))
    for f in system.forces:
        t = f.type
        if t == "HarmonicBondForce" and f.count:
            d.update(n_bonds=f.count, bond_p1=i32(f.p1), bond_p2=i32(f.p2), bond_length=f64(f.length),
                     bond_k=f64(f.k))
        elif t == "HarmonicAngleForce" and f.count:
            d.update(n_angles=f.count, angle_p1=i32(f.p1), angle_p2=i32(f.p2), angle_p3=i32(f.p3),
                     angle_theta=f64(f.angle), angle_k=f64(f.k))
        elif t == "PeriodicTorsionForce" and f.count:
            d.update(n_torsions=f.count, torsion_p1=i32(f.p1), torsion_p2=i32(f.p2), torsion_p3=i32(f.p3),
                     torsion_p4=i32(f.p4), torsion_periodicity=i32(f.periodicity), torsion_phase=f64(f.phase),
                     torsion_k=f64(f.k))
        elif t == "NonbondedForce":
            if f.method != 4:
                raise MdError(1, f"NonbondedForce method {f.method}: this engine implements PME (4)")
            if f.useSwitchingFunction:
                raise MdError(1, "NonbondedForce with a switching function: this engine implements the plain "
                                 "cutoff of the reference pair term")
            d.update(has_nonbonded=1, cutoff=float(f.cutoff), ewald_tolerance=float(f.ewaldTolerance),
                     alpha=float(f.alpha), grid=i32([int(g) for g in f.grid]),
                     use_dispersion_correction=int(bool(f.useDispersionCorrection)),
                     charge=f64(f.charge), sigma=f64(f.sigma), epsilon=f64(f.epsilon),
                     n_exceptions=f.exceptionCount)
            if f.exceptionCount:
                d.update(exception_p1=i32(f.exceptionP1), exception_p2=i32(f.exceptionP2),
                         exception_chargeprod=f64(f.exceptionChargeProd), exception_sigma=f64(f.exceptionSigma),
                         exception_epsilon=f64(f.exceptionEpsilon))
        elif t == "PositionRestraint" and f.count:
            d.update(n_restraints=f.count, restraint_particle=i32(f.particle), restraint_k=f64(f.k),
                     restraint_x0=f64(f.x0), restraint_y0=f64(f.y0), restraint_z0=f64(f.z0))
        elif t == "CMMotionRemover":
            d["cmm_frequency"] = int(f.frequency)
        elif t == "MonteCarloBarostat":
            d.update(has_barostat=1, barostat_pressure=float(f.defaultPressure),
                     barostat_temperature=float(f.defaultTemperature), barostat_frequency=int(f.frequency),
                     barostat_seed=int(f.randomSeed) & 0xFFFFFFFF)
    return d


def describe(system, box: Optional[Sequence] = None):
    """Pack an OpenMMSystem into ``md_system_desc``.  Returns (desc, keepalive)."""
    fields = describe_fields(system, box)
    keep = []
    d = SystemDesc()
    for name, value in fields.items():
        if name in ("box", "grid"):
            arr = getattr(d, name)
            for k in range(len(arr)):
                arr[k] = value[k]
        elif isinstance(value, np.ndarray):
            keep.append(value)
            setattr(d, name, _d(value) if value.dtype == np.float64 else _i(value))
        else:
            setattr(d, name, value)
    return d, keep


class NativeSystem:
    """An ``md_system*`` handle: one resident system on one GPU."""

    def __init__(self, system, time_step=0.002, temperature=300.0, friction=1.0,
                 constraint_tolerance=1e-6, seed=0, integrator="LangevinMiddle", verlet_skin=-1.0,
                 mesh=(0, 0, 0), device=0, box=None, max_batch=0):
        L = library()
        desc, keep = describe(system, box)
        o = RunOptionsC()
        o.time_step, o.temperature, o.friction = time_step, temperature, friction
        o.constraint_tolerance = constraint_tolerance
        o.seed = int(seed) & 0xFFFFFFFF
        o.integrator = 0 if integrator == "LangevinMiddle" else 1
        o.verlet_skin = verlet_skin
        for k in range(3):
            o.mesh[k] = int(mesh[k])
        o.device = device
        o.max_batch = int(max_batch)
        h = C.c_void_p()
        check(L.md_system_create(C.byref(desc), C.byref(o), C.byref(h)))
        self._h = h
        self.n = system.particleCount

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            library().md_system_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("the system has been destroyed")
        return self._h

    def set_positions(self, x):
        a = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        check(library().md_set_positions(self.handle, _d(a)))

    def set_velocities(self, v):
        a = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        check(library().md_set_velocities(self.handle, _d(a)))

    def positions(self) -> np.ndarray:
        out = np.zeros(3 * self.n)
        check(library().md_get_positions(self.handle, _d(out)))
        return out

    def velocities(self) -> np.ndarray:
        out = np.zeros(3 * self.n)
        check(library().md_get_velocities(self.handle, _d(out)))
        return out

    def set_box(self, box):
        b = np.ascontiguousarray(np.asarray(box, dtype=np.float64).reshape(9))
        check(library().md_set_box(self.handle, _d(b)))

    def box(self) -> np.ndarray:
        b = np.zeros(9)
        check(library().md_get_box(self.handle, _d(b)))
        return b.reshape(3, 3)

    def mesh(self):
        m = np.zeros(3, dtype=np.int32)
        check(library().md_get_mesh(self.handle, _i(m)))
        return tuple(int(v) for v in m)

    def evaluate(self, forces=True):
        """(total, groups[6], forces[3N] or None)."""
        f = np.zeros(3 * self.n) if forces else None
        g = np.zeros(6)
        t = C.c_double()
        check(library().md_evaluate(self.handle, _d(f) if forces else None, _d(g), C.byref(t)))
        return t.value, g, f

    def run(self, steps: int):
        check(library().md_run(self.handle, int(steps)))

    @property
    def step(self) -> int:
        v = C.c_int64()
        check(library().md_get_step(self.handle, C.byref(v)))
        return v.value

    @step.setter
    def step(self, value: int):
        check(library().md_set_step(self.handle, int(value)))

    def list_flags(self) -> dict:
        f = ListFlags()
        check(library().md_list_flags_get(self.handle, C.byref(f)))
        return {"rebuilds": f.rebuilds, "tiles": f.tiles, "tileCapacity": f.tile_capacity,
                "worstDisplacement": f.worst_displacement, "overflowed": bool(f.overflowed)}

    def pair_census(self):
        p, t = C.c_int64(), C.c_int64()
        check(library().md_pair_census(self.handle, C.byref(p), C.byref(t)))
        return p.value, t.value

    def inject_fault(self, after: int):
        check(library().md_inject_fault(self.handle, int(after)))

    def profile_steps(self, steps: int) -> dict:
        """Mean milliseconds per step of each phase, timed with HIP events."""
        ms = np.zeros(PHASE_COUNT)
        check(library().md_profile_steps(self.handle, int(steps), _d(ms)))
        return {library().md_phase_name(k).decode(): float(ms[k]) for k in range(PHASE_COUNT)}

    def autotune(self, steps: int) -> dict:
        out = Schedule()
        check(library().md_autotune(self.handle, int(steps), C.byref(out)))
        return out.as_dict()

    def set_pair_schedule(self, segment_entries: int = 0, pair_waves: int = 0, compact: bool = False):
        """md_set_pair_schedule (0 keeps the current length or wave count)."""
        sc = Schedule(segment_entries=int(segment_entries), pair_waves=int(pair_waves), compact_pairs=int(bool(compact)))
        check(library().md_set_pair_schedule(self.handle, C.byref(sc)))

    def minimise(self, max_iterations: int = 0, tolerance: float = 10.0) -> dict:
        """md_minimise: L-BFGS with constraint springs, searched on the device."""
        out = MinimiseResultC()
        check(library().md_minimise(self.handle, int(max_iterations), float(tolerance), C.byref(out)))
        return {k: getattr(out, k) for k, _ in out._fields_}

    def attach_barostat(self, pressure: float, temperature: float, frequency: int, molecule_start,
                        molecule_atoms, seed: int = 0, volume_scale: float = 0.0, draws: int = 0,
                        steps_since_attempt: int = 0, window_attempted: int = 0, window_accepted: int = 0):
        """md_barostat_attach: the Monte Carlo barostat run on the device inside
        ``run``; ``frequency`` 0 detaches it."""
        start = np.ascontiguousarray(molecule_start, dtype=np.int32)
        atoms = np.ascontiguousarray(molecule_atoms, dtype=np.int32)
        d = BarostatDescC(float(pressure), float(temperature), int(frequency), int(start.shape[0] - 1),
                          start.ctypes.data_as(_ip), atoms.ctypes.data_as(_ip), int(seed) & 0xFFFFFFFF,
                          float(volume_scale), int(draws), int(steps_since_attempt), int(window_attempted),
                          int(window_accepted))
        check(library().md_barostat_attach(self.handle, C.byref(d)))

    def barostat_state(self) -> dict:
        out = BarostatStateC()
        check(library().md_barostat_get(self.handle, C.byref(out)))
        return {k: getattr(out, k) for k, _ in out._fields_}

    def constrain_positions(self, tolerance: float):
        check(library().md_constrain_positions(self.handle, float(tolerance)))

    def constrain_velocities(self, tolerance: float):
        check(library().md_constrain_velocities(self.handle, float(tolerance)))


class NativeBatch:
    """An ``md_batch*``: several ``NativeSystem`` replicas advanced together,
    one kernel launch per phase for all of them.  Holds references to its
    systems so they outlive it."""

    def __init__(self, systems):
        self.systems = list(systems)
        arr = (C.c_void_p * len(self.systems))(*[s.handle.value for s in self.systems])
        h = C.c_void_p()
        check(library().md_batch_create(arr, len(self.systems), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            library().md_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, steps: int):
        check(library().md_batch_run(self._h, int(steps)))

    def profile_steps(self, steps: int) -> dict:
        ms = np.zeros(PHASE_COUNT)
        check(library().md_batch_profile_steps(self._h, int(steps), _d(ms)))
        return {library().md_phase_name(k).decode(): float(ms[k]) for k in range(PHASE_COUNT)}

    def autotune(self, steps: int) -> dict:
        """md_batch_autotune: the fastest pair-kernel schedule, and the time of
        the one the batch started with."""
        out = Schedule()
        check(library().md_batch_autotune(self._h, int(steps), C.byref(out)))
        return out.as_dict()

    def set_pair_schedule(self, segment_entries: int = 0, pair_waves: int = 0, compact: bool = False):
        """md_batch_set_pair_schedule (0 keeps the current length or wave count)."""
        sc = Schedule(segment_entries=int(segment_entries), pair_waves=int(pair_waves), compact_pairs=int(bool(compact)))
        check(library().md_batch_set_pair_schedule(self._h, C.byref(sc)))


PHASE_COUNT = 12


def probe_fixed_point(values):
    v = np.ascontiguousarray(values, dtype=np.float32)
    fixed = np.zeros(v.shape[0], dtype=np.int64)
    back = np.zeros(v.shape[0], dtype=np.float32)
    check(library().md_probe_fixed_point(v.ctypes.data_as(C.POINTER(C.c_float)), v.shape[0],
                                         fixed.ctypes.data_as(C.POINTER(C.c_int64)),
                                         back.ctypes.data_as(C.POINTER(C.c_float))))
    return fixed, back


def probe_cell_order(dims):
    """The neighbour search's cell sort order for a grid of ``dims`` cells
    (host only)."""
    d = np.asarray(dims, dtype=np.int32)
    cells = np.zeros(int(np.prod(d)), dtype=np.int32)
    check(library().md_probe_cell_order(_i(d), _i(cells)))
    return cells


def probe_ewald_fit(alpha: float, cutoff: float):
    """The pair kernel's Coulomb polynomials for ``alpha`` (1/nm) and
    ``cutoff`` (nm), host only: a dict with ``scale``, the 13 f32 ``force``
    and ``energy`` coefficients, ``ok`` and the fit's ``force_err`` and
    ``energy_err`` (relative to 1/r^3 and 1/r)."""
    scale = C.c_float()
    force = np.zeros(13, dtype=np.float32)
    energy = np.zeros(13, dtype=np.float32)
    ok = C.c_int32()
    ferr, eerr = C.c_double(), C.c_double()
    fp = C.POINTER(C.c_float)
    check(library().md_probe_ewald_fit(alpha, cutoff, C.byref(scale), force.ctypes.data_as(fp),
                                       energy.ctypes.data_as(fp), C.byref(ok), C.byref(ferr), C.byref(eerr)))
    return {"scale": np.float32(scale.value), "force": force, "energy": energy, "ok": bool(ok.value),
            "force_err": ferr.value, "energy_err": eerr.value}


def probe_thermostat(seed: int, step: int, n: int):
    bits = np.zeros(4 * n, dtype=np.uint32)
    normals = np.zeros(3 * n, dtype=np.float32)
    check(library().md_probe_thermostat(seed, step, n, bits.ctypes.data_as(C.POINTER(C.c_uint32)),
                                        normals.ctypes.data_as(C.POINTER(C.c_float))))
    return bits.reshape(n, 4), normals.reshape(n, 3)


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


class NativeWarmStart:
    """An ``md_warmstart*``: the learned SHAKE warm-start network, its
    recorded training samples and its optimiser state, on one device."""

    F, MAXC, O = 32, 4, 16

    def __init__(self, device=0, seed=0, capacity=1_000_000):
        h = C.c_void_p()
        check(library().md_warmstart_create(int(device), int(seed) & 0xFFFFFFFF, int(capacity), C.byref(h)))
        self._h = h
        self.nparams = int(library().md_warmstart_param_count())

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            library().md_warmstart_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def record(self, system: "NativeSystem", steps: int):
        check(library().md_warmstart_record(self._h, system.handle, int(steps)))

    @property
    def samples(self) -> int:
        v = C.c_int64()
        check(library().md_warmstart_sample_count(self._h, C.byref(v)))
        return v.value

    def train(self, iterations: int, batch: int = 4096, learning_rate: float = 1e-3) -> float:
        loss = C.c_double()
        check(library().md_warmstart_train(self._h, int(iterations), int(batch), float(learning_rate),
                                           C.byref(loss)))
        return loss.value

    def attach(self, system: "NativeSystem"):
        check(library().md_warmstart_attach(system.handle, self._h))

    @staticmethod
    def detach(system: "NativeSystem"):
        check(library().md_warmstart_attach(system.handle, None))

    def params(self) -> np.ndarray:
        out = np.zeros(self.nparams, dtype=np.float32)
        check(library().md_warmstart_params(self._h, _f(out), None))
        return out

    def set_params(self, theta):
        t = np.ascontiguousarray(theta, dtype=np.float32)
        check(library().md_warmstart_params(self._h, None, _f(t)))

    def evaluate(self, x, t, outputs=True, gradient=True):
        """(loss, outputs [n, 16] or None, gradient [nparams] or None)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        n = x.shape[0]
        y = np.zeros((n, self.O), dtype=np.float32) if outputs else None
        g = np.zeros(self.nparams, dtype=np.float32) if gradient else None
        loss = C.c_double()
        check(library().md_warmstart_evaluate(self._h, _f(x), _f(t), n, _f(y), _f(g), C.byref(loss)))
        return loss.value, y, g


def shake_counts(system: "NativeSystem"):
    """(SHAKE update passes, cluster-steps) of a system so far."""
    p, c = C.c_int64(), C.c_int64()
    check(library().md_shake_counts(system.handle, C.byref(p), C.byref(c)))
    return p.value, c.value
