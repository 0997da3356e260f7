"""Streaming System / State ingestion through libmdcore's C reader.

``read_system`` and ``read_state`` give the same ``OpenMMSystem`` /
``OpenMMState`` as ``parseSystem`` / ``parseState`` (``system.py``,
``state.py``), with the same refusals, but the document is read in one pass
by ``md_xml_parse_system`` / ``md_xml_parse_state`` (no element tree), so a
250k-atom System parses in about a second instead of the tens of seconds an
element tree needs (``src/engine/system.ts:515-520`` notes the scale upstream).
Force groups and names are not carried (the engine does not use them); the
forces come out in a fixed order: bonds, angles, torsions, nonbonded,
restraint, COM remover, barostat.
"""
from __future__ import annotations

import ctypes as C
from typing import Union

import numpy as np

from . import native
from .state import OpenMMState
from .system import (CMMotionRemover, HarmonicAngleForce, HarmonicBondForce, MonteCarloBarostat, NonbondedForce,
                     OpenMMSystem, PeriodicTorsionForce, PositionRestraintForce, UnsupportedForce, XmlError)


def _bind():
    L = native.library()
    if not getattr(L, "_ingest_bound", False):
        L.md_xml_parse_system.argtypes = [C.c_char_p, C.c_int64, C.POINTER(C.c_void_p)]
        L.md_xml_system_describe.argtypes = [C.c_void_p, C.POINTER(native.SystemDesc)]
        L.md_xml_system_free.argtypes = [C.c_void_p]
        L.md_xml_system_free.restype = None
        L.md_xml_parse_state.argtypes = [C.c_char_p, C.c_int64, C.POINTER(C.c_void_p)]
        L.md_xml_state_get.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.POINTER(C.c_double)),
                                       C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.c_double),
                                       C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        L.md_xml_state_free.argtypes = [C.c_void_p]
        L.md_xml_state_free.restype = None
        L._ingest_bound = True
    return L


def _raise(code: int):
    msg = native.library().md_last_error().decode(errors="replace")
    if code == 1:
        raise UnsupportedForceText(msg)
    raise XmlError(msg)


class UnsupportedForceText(UnsupportedForce):
    """UnsupportedForce carrying the C reader's message as it is."""

    def __init__(self, message: str):
        ValueError.__init__(self, message)
        self.forceType = message.split(" ")[0].rstrip(":")


def _text(source: Union[str, bytes]) -> bytes:
    if isinstance(source, bytes):
        return source
    if source.lstrip().startswith("<"):
        return source.encode()
    with open(source, "rb") as f:
        return f.read()


def _f(ptr, n):
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(np.float64, copy=True) if n else np.zeros(0)


def _i(ptr, n):
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(np.int32, copy=True) if n else np.zeros(0, np.int32)


def read_system(source: Union[str, bytes]) -> OpenMMSystem:
    """An OpenMM System from XML text, bytes or a file path."""
    L = _bind()
    data = _text(source)
    h = C.c_void_p()
    rc = L.md_xml_parse_system(data, len(data), C.byref(h))
    if rc:
        _raise(rc)
    try:
        d = native.SystemDesc()
        native.check(L.md_xml_system_describe(h, C.byref(d)))
        n = d.n_particles
        box = np.array(d.box[:], dtype=np.float64).reshape(3, 3)
        forces = []
        if d.n_bonds:
            forces.append(HarmonicBondForce(_i(d.bond_p1, d.n_bonds), _i(d.bond_p2, d.n_bonds),
                                            _f(d.bond_length, d.n_bonds), _f(d.bond_k, d.n_bonds)))
        if d.n_angles:
            forces.append(HarmonicAngleForce(_i(d.angle_p1, d.n_angles), _i(d.angle_p2, d.n_angles),
                                             _i(d.angle_p3, d.n_angles), _f(d.angle_theta, d.n_angles),
                                             _f(d.angle_k, d.n_angles)))
        if d.n_torsions:
            t = d.n_torsions
            forces.append(PeriodicTorsionForce(_i(d.torsion_p1, t), _i(d.torsion_p2, t), _i(d.torsion_p3, t),
                                               _i(d.torsion_p4, t), _i(d.torsion_periodicity, t),
                                               _f(d.torsion_phase, t), _f(d.torsion_k, t)))
        if d.has_nonbonded:
            e = d.n_exceptions
            forces.append(NonbondedForce(
                charge=_f(d.charge, n), sigma=_f(d.sigma, n), epsilon=_f(d.epsilon, n),
                exceptionP1=_i(d.exception_p1, e), exceptionP2=_i(d.exception_p2, e),
                exceptionChargeProd=_f(d.exception_chargeprod, e), exceptionSigma=_f(d.exception_sigma, e),
                exceptionEpsilon=_f(d.exception_epsilon, e), cutoff=d.cutoff, ewaldTolerance=d.ewald_tolerance,
                alpha=d.alpha, grid=tuple(int(g) for g in d.grid),
                useDispersionCorrection=bool(d.use_dispersion_correction)))
        if d.n_restraints:
            r = d.n_restraints
            forces.append(PositionRestraintForce(_i(d.restraint_particle, r), _f(d.restraint_k, r),
                                                 _f(d.restraint_x0, r), _f(d.restraint_y0, r),
                                                 _f(d.restraint_z0, r)))
        if d.cmm_frequency:
            forces.append(CMMotionRemover(int(d.cmm_frequency)))
        if d.has_barostat:
            forces.append(MonteCarloBarostat(d.barostat_pressure, d.barostat_temperature, int(d.barostat_frequency),
                                             int(d.barostat_seed)))
        c = d.n_constraints
        return OpenMMSystem(boxVectors=tuple(tuple(float(v) for v in row) for row in box),
                            masses=_f(d.masses, n), constraintP1=_i(d.constraint_p1, c),
                            constraintP2=_i(d.constraint_p2, c), constraintDistance=_f(d.constraint_distance, c),
                            forces=forces)
    finally:
        L.md_xml_system_free(h)


def read_state(source: Union[str, bytes]) -> OpenMMState:
    """An OpenMM State from XML text, bytes or a file path."""
    L = _bind()
    data = _text(source)
    h = C.c_void_p()
    rc = L.md_xml_parse_state(data, len(data), C.byref(h))
    if rc:
        _raise(rc)
    try:
        n = C.c_int32()
        px, pv = C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
        box = (C.c_double * 9)()
        step, time = C.c_int64(), C.c_double()
        native.check(L.md_xml_state_get(h, C.byref(n), C.byref(px), C.byref(pv), box, C.byref(step), C.byref(time)))
        b = np.array(box[:], dtype=np.float64).reshape(3, 3)
        return OpenMMState(boxVectors=tuple(tuple(float(v) for v in row) for row in b),
                           positions=_f(px, 3 * n.value),
                           velocities=_f(pv, 3 * n.value) if pv else None,
                           stepCount=int(step.value), time=float(time.value))
    finally:
        L.md_xml_state_free(h)
