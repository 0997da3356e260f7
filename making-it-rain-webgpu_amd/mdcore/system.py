"""The engine's input contract: an OpenMM ``System`` as ``XmlSerializer`` writes it.

Mirrors ``src/engine/system.ts`` of the upstream project: the same seven force
kinds are accepted (``system.ts:20-173``), anything else raises
``UnsupportedForce`` (``system.ts:178-190``), parameter offsets and LJ-PME are
refused (``system.ts:371-380``), and a ``CustomExternalForce`` is accepted only
for the restraint expression ``k*periodicdistance(x,y,z,x0,y0,z0)^2``
(``system.ts:402-412``).  Arrays are numpy: ``int32`` indices and ``float64``
parameters, positions xyz-interleaved.
"""
from __future__ import annotations

import math
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple, Union

import numpy as np

Vec3 = Tuple[float, float, float]
BoxVectors = Tuple[Vec3, Vec3, Vec3]

RESTRAINT_EXPRESSION = "k*periodicdistance(x,y,z,x0,y0,z0)^2"


class NonbondedMethod:
    """OpenMM's ``NonbondedForce::NonbondedMethod`` (``system.ts:23-30``)."""

    NoCutoff = 0
    CutoffNonPeriodic = 1
    CutoffPeriodic = 2
    Ewald = 3
    PME = 4
    LJPME = 5


class XmlError(ValueError):
    """A malformed document (``src/engine/xml.ts``)."""


class UnsupportedForce(ValueError):
    """A force this engine does not implement (``system.ts:178-190``)."""

    def __init__(self, force_type: str, detail: Optional[str] = None):
        self.forceType = force_type
        message = f"{force_type} is not implemented by this engine"
        if detail is not None:
            message += f": {detail}"
        message += (
            ". The supported set is HarmonicBondForce, HarmonicAngleForce, "
            "PeriodicTorsionForce, NonbondedForce, CMMotionRemover, "
            'MonteCarloBarostat, and a CustomExternalForce holding exactly '
            f'"{RESTRAINT_EXPRESSION}".'
        )
        super().__init__(message)


@dataclass
class HarmonicBondForce:
    """E = 1/2 k (r - length)^2 (``system.ts:40-50``)."""

    p1: np.ndarray
    p2: np.ndarray
    length: np.ndarray
    k: np.ndarray
    forceGroup: int = 0
    name: str = "HarmonicBondForce"
    usesPeriodic: bool = False
    type: str = "HarmonicBondForce"

    @property
    def count(self) -> int:
        return int(self.p1.shape[0])


@dataclass
class HarmonicAngleForce:
    """E = 1/2 k (theta - angle)^2 (``system.ts:53-65``)."""

    p1: np.ndarray
    p2: np.ndarray
    p3: np.ndarray
    angle: np.ndarray
    k: np.ndarray
    forceGroup: int = 0
    name: str = "HarmonicAngleForce"
    usesPeriodic: bool = False
    type: str = "HarmonicAngleForce"

    @property
    def count(self) -> int:
        return int(self.p1.shape[0])


@dataclass
class PeriodicTorsionForce:
    """E = k (1 + cos(n phi - phase)) (``system.ts:68-82``)."""

    p1: np.ndarray
    p2: np.ndarray
    p3: np.ndarray
    p4: np.ndarray
    periodicity: np.ndarray
    phase: np.ndarray
    k: np.ndarray
    forceGroup: int = 0
    name: str = "PeriodicTorsionForce"
    usesPeriodic: bool = False
    type: str = "PeriodicTorsionForce"

    @property
    def count(self) -> int:
        return int(self.p1.shape[0])


@dataclass
class NonbondedForce:
    """PME NonbondedForce (``system.ts:84-115``)."""

    charge: np.ndarray
    sigma: np.ndarray
    epsilon: np.ndarray
    exceptionP1: np.ndarray
    exceptionP2: np.ndarray
    exceptionChargeProd: np.ndarray
    exceptionSigma: np.ndarray
    exceptionEpsilon: np.ndarray
    cutoff: float = 1.0
    ewaldTolerance: float = 5e-4
    alpha: float = 0.0
    grid: Vec3 = (0, 0, 0)
    method: int = NonbondedMethod.PME
    useDispersionCorrection: bool = True
    useSwitchingFunction: bool = False
    switchingDistance: float = -1.0
    exceptionsUsePeriodic: bool = False
    includeDirectSpace: bool = True
    recipForceGroup: int = -1
    forceGroup: int = 0
    name: str = "NonbondedForce"
    type: str = "NonbondedForce"

    @property
    def particleCount(self) -> int:
        return int(self.charge.shape[0])

    @property
    def exceptionCount(self) -> int:
        return int(self.exceptionP1.shape[0])


@dataclass
class CMMotionRemover:
    frequency: int = 1
    forceGroup: int = 0
    name: str = "CMMotionRemover"
    type: str = "CMMotionRemover"


@dataclass
class MonteCarloBarostat:
    defaultPressure: float = 1.0
    defaultTemperature: float = 300.0
    frequency: int = 25
    randomSeed: int = 0
    forceGroup: int = 0
    name: str = "MonteCarloBarostat"
    type: str = "MonteCarloBarostat"


@dataclass
class PositionRestraintForce:
    """``k*periodicdistance(x,y,z,x0,y0,z0)^2`` (``system.ts:128-143``)."""

    particle: np.ndarray
    k: np.ndarray
    x0: np.ndarray
    y0: np.ndarray
    z0: np.ndarray
    energyExpression: str = RESTRAINT_EXPRESSION
    forceGroup: int = 0
    name: str = "CustomExternalForce"
    type: str = "PositionRestraint"

    @property
    def count(self) -> int:
        return int(self.particle.shape[0])


Force = Union[
    HarmonicBondForce,
    HarmonicAngleForce,
    PeriodicTorsionForce,
    NonbondedForce,
    CMMotionRemover,
    MonteCarloBarostat,
    PositionRestraintForce,
]


@dataclass
class OpenMMSystem:
    """``OpenMMSystem`` (``system.ts:161-173``)."""

    boxVectors: BoxVectors
    masses: np.ndarray
    constraintP1: np.ndarray
    constraintP2: np.ndarray
    constraintDistance: np.ndarray
    forces: List[Force] = field(default_factory=list)
    openmmVersion: str = "8.4"

    @property
    def particleCount(self) -> int:
        return int(self.masses.shape[0])

    @property
    def constraintCount(self) -> int:
        return int(self.constraintP1.shape[0])

    def force(self, type_name: str):
        for f in self.forces:
            if f.type == type_name:
                return f
        return None


def _i32(values: Sequence[int]) -> np.ndarray:
    return np.asarray(values, dtype=np.int32)


def _f64(values: Sequence[float]) -> np.ndarray:
    return np.asarray(values, dtype=np.float64)


def _num(el: ET.Element, name: str, default: Optional[float] = None) -> float:
    value = el.get(name)
    if value is None:
        if default is None:
            raise XmlError(f"<{el.tag}> has no attribute {name}")
        return default
    return float(value)


def _flag(el: ET.Element, name: str, default: bool = False) -> bool:
    value = el.get(name)
    if value is None:
        return default
    return value not in ("0", "false", "False")


def _box(el: Optional[ET.Element]) -> BoxVectors:
    if el is None:
        raise XmlError("no <PeriodicBoxVectors>")
    rows = []
    for tag in ("A", "B", "C"):
        v = el.find(tag)
        if v is None:
            raise XmlError(f"expected three box vectors, missing {tag}")
        rows.append((_num(v, "x"), _num(v, "y"), _num(v, "z")))
    return (rows[0], rows[1], rows[2])


def _read_force(el: ET.Element) -> Force:
    kind = el.get("type", "")
    group = int(_num(el, "forceGroup", 0))
    name = el.get("name", kind)
    periodic = _flag(el, "usesPeriodic")
    if kind == "HarmonicBondForce":
        bonds = el.findall("./Bonds/Bond")
        return HarmonicBondForce(
            _i32([int(b.get("p1")) for b in bonds]),
            _i32([int(b.get("p2")) for b in bonds]),
            _f64([_num(b, "d") for b in bonds]),
            _f64([_num(b, "k") for b in bonds]),
            group, name, periodic)
    if kind == "HarmonicAngleForce":
        angles = el.findall("./Angles/Angle")
        return HarmonicAngleForce(
            _i32([int(a.get("p1")) for a in angles]),
            _i32([int(a.get("p2")) for a in angles]),
            _i32([int(a.get("p3")) for a in angles]),
            _f64([_num(a, "a") for a in angles]),
            _f64([_num(a, "k") for a in angles]),
            group, name, periodic)
    if kind == "PeriodicTorsionForce":
        ts = el.findall("./Torsions/Torsion")
        return PeriodicTorsionForce(
            _i32([int(t.get("p1")) for t in ts]),
            _i32([int(t.get("p2")) for t in ts]),
            _i32([int(t.get("p3")) for t in ts]),
            _i32([int(t.get("p4")) for t in ts]),
            _i32([int(t.get("periodicity")) for t in ts]),
            _f64([_num(t, "phase") for t in ts]),
            _f64([_num(t, "k") for t in ts]),
            group, name, periodic)
    if kind == "NonbondedForce":
        offsets = len(el.findall("./ParticleOffsets/*")) + len(el.findall("./ExceptionOffsets/*"))
        if offsets > 0:
            raise UnsupportedForce(
                "NonbondedForce",
                f"it carries {offsets} parameter offset(s), which make the energy "
                "depend on a global parameter")
        method = int(_num(el, "method"))
        if method == NonbondedMethod.LJPME:
            raise UnsupportedForce("NonbondedForce", "LJ-PME is not implemented")
        ps = el.findall("./Particles/Particle")
        es = el.findall("./Exceptions/Exception")
        return NonbondedForce(
            charge=_f64([_num(p, "q") for p in ps]),
            sigma=_f64([_num(p, "sig") for p in ps]),
            epsilon=_f64([_num(p, "eps") for p in ps]),
            exceptionP1=_i32([int(e.get("p1")) for e in es]),
            exceptionP2=_i32([int(e.get("p2")) for e in es]),
            exceptionChargeProd=_f64([_num(e, "q") for e in es]),
            exceptionSigma=_f64([_num(e, "sig") for e in es]),
            exceptionEpsilon=_f64([_num(e, "eps") for e in es]),
            cutoff=_num(el, "cutoff"),
            ewaldTolerance=_num(el, "ewaldTolerance"),
            alpha=_num(el, "alpha", 0.0),
            grid=(int(_num(el, "nx", 0)), int(_num(el, "ny", 0)), int(_num(el, "nz", 0))),
            method=method,
            useDispersionCorrection=_flag(el, "dispersionCorrection"),
            useSwitchingFunction=_flag(el, "useSwitchingFunction"),
            switchingDistance=_num(el, "switchingDistance", -1.0),
            exceptionsUsePeriodic=_flag(el, "exceptionsUsePeriodic"),
            includeDirectSpace=_flag(el, "includeDirectSpace", True),
            recipForceGroup=int(_num(el, "recipForceGroup", -1)),
            forceGroup=group, name=name)
    if kind == "CMMotionRemover":
        return CMMotionRemover(int(_num(el, "frequency", 1)), group, name)
    if kind == "MonteCarloBarostat":
        return MonteCarloBarostat(_num(el, "pressure"), _num(el, "temperature"),
                                  int(_num(el, "frequency", 25)), int(_num(el, "randomSeed", 0)),
                                  group, name)
    if kind == "CustomExternalForce":
        expression = el.get("energy", "")
        normalised = "".join(ch for ch in expression if not ch.isspace() and ch != ";")
        if normalised != RESTRAINT_EXPRESSION:
            raise UnsupportedForce(
                "CustomExternalForce",
                f'its energy expression is "{expression}", and this engine implements '
                f'only "{RESTRAINT_EXPRESSION}"')
        globals_ = [p.get("name") for p in el.findall("./GlobalParameters/Parameter")]
        if globals_:
            raise UnsupportedForce(
                "CustomExternalForce",
                f"it has global parameters ({', '.join(globals_)})")
        names = [p.get("name") for p in el.findall("./PerParticleParameters/Parameter")]
        rows = []
        idx = []
        for p in el.findall("./Particles/Particle"):
            idx.append(int(p.get("index")))
            rows.append([_num(p, f"param{i + 1}") for i in range(len(names))])

        def column(want: str) -> np.ndarray:
            if want not in names:
                raise UnsupportedForce(
                    "CustomExternalForce",
                    f'it has no per-particle parameter "{want}" (found {", ".join(names)})')
            c = names.index(want)
            return _f64([r[c] for r in rows])

        return PositionRestraintForce(_i32(idx), column("k"), column("x0"), column("y0"),
                                      column("z0"), expression, group, name)
    raise UnsupportedForce(kind)


def parseSystem(xml: str) -> OpenMMSystem:
    """Parse System XML (``system.ts:521-597``)."""
    try:
        root = ET.fromstring(xml)
    except ET.ParseError as error:
        raise XmlError(str(error)) from error
    if root.tag != "System":
        raise XmlError("no <System> element")
    box = _box(root.find("PeriodicBoxVectors"))
    masses = _f64([_num(p, "mass") for p in root.findall("./Particles/Particle")])
    cons = root.findall("./Constraints/Constraint")
    forces = [_read_force(f) for f in root.findall("./Forces/Force")]
    system = OpenMMSystem(
        boxVectors=box,
        masses=masses,
        constraintP1=_i32([int(c.get("p1")) for c in cons]),
        constraintP2=_i32([int(c.get("p2")) for c in cons]),
        constraintDistance=_f64([_num(c, "d") for c in cons]),
        forces=forces,
        openmmVersion=root.get("openmmVersion", ""),
    )
    nb = system.force("NonbondedForce")
    if nb is not None and nb.particleCount != system.particleCount:
        raise XmlError(
            f"NonbondedForce has {nb.particleCount} particles but the System has "
            f"{system.particleCount}")
    return system


def _n(value: float) -> str:
    return repr(float(value))


def serialiseSystem(system: OpenMMSystem) -> str:
    """Write System XML readable by OpenMM (``system.ts:619-``)."""
    out = ['<?xml version="1.0" ?>',
           f'<System openmmVersion="{system.openmmVersion}" type="System" version="1">',
           "\t<PeriodicBoxVectors>"]
    for label, v in zip("ABC", system.boxVectors):
        out.append(f'\t\t<{label} x="{_n(v[0])}" y="{_n(v[1])}" z="{_n(v[2])}"/>')
    out.append("\t</PeriodicBoxVectors>")
    out.append("\t<Particles>")
    out.extend(f'\t\t<Particle mass="{_n(m)}"/>' for m in system.masses)
    out.append("\t</Particles>")
    if system.constraintCount == 0:
        out.append("\t<Constraints/>")
    else:
        out.append("\t<Constraints>")
        for a, b, d in zip(system.constraintP1, system.constraintP2, system.constraintDistance):
            out.append(f'\t\t<Constraint d="{_n(d)}" p1="{int(a)}" p2="{int(b)}"/>')
        out.append("\t</Constraints>")
    out.append("\t<Forces>")
    for f in system.forces:
        common = f'forceGroup="{f.forceGroup}" name="{f.name}"'
        if f.type == "HarmonicBondForce":
            out.append(f'\t\t<Force {common} type="HarmonicBondForce" usesPeriodic="0" version="2">')
            out.append("\t\t\t<Bonds>")
            for i in range(f.count):
                out.append(f'\t\t\t\t<Bond d="{_n(f.length[i])}" k="{_n(f.k[i])}" '
                           f'p1="{int(f.p1[i])}" p2="{int(f.p2[i])}"/>')
            out.append("\t\t\t</Bonds>")
        elif f.type == "HarmonicAngleForce":
            out.append(f'\t\t<Force {common} type="HarmonicAngleForce" usesPeriodic="0" version="2">')
            out.append("\t\t\t<Angles>")
            for i in range(f.count):
                out.append(f'\t\t\t\t<Angle a="{_n(f.angle[i])}" k="{_n(f.k[i])}" '
                           f'p1="{int(f.p1[i])}" p2="{int(f.p2[i])}" p3="{int(f.p3[i])}"/>')
            out.append("\t\t\t</Angles>")
        elif f.type == "PeriodicTorsionForce":
            out.append(f'\t\t<Force {common} type="PeriodicTorsionForce" usesPeriodic="0" version="2">')
            out.append("\t\t\t<Torsions>")
            for i in range(f.count):
                out.append(f'\t\t\t\t<Torsion k="{_n(f.k[i])}" p1="{int(f.p1[i])}" p2="{int(f.p2[i])}" '
                           f'p3="{int(f.p3[i])}" p4="{int(f.p4[i])}" '
                           f'periodicity="{int(f.periodicity[i])}" phase="{_n(f.phase[i])}"/>')
            out.append("\t\t\t</Torsions>")
        elif f.type == "NonbondedForce":
            out.append(
                f'\t\t<Force alpha="{_n(f.alpha)}" cutoff="{_n(f.cutoff)}" '
                f'dispersionCorrection="{int(f.useDispersionCorrection)}" '
                f'ewaldTolerance="{_n(f.ewaldTolerance)}" '
                f'exceptionsUsePeriodic="{int(f.exceptionsUsePeriodic)}" {common} '
                f'includeDirectSpace="{int(f.includeDirectSpace)}" method="{f.method}" '
                f'nx="{int(f.grid[0])}" ny="{int(f.grid[1])}" nz="{int(f.grid[2])}" '
                f'recipForceGroup="{f.recipForceGroup}" rfDielectric="78.3" '
                f'switchingDistance="{_n(f.switchingDistance)}" type="NonbondedForce" '
                f'useSwitchingFunction="{int(f.useSwitchingFunction)}" version="4">')
            out.append("\t\t\t<GlobalParameters/>\n\t\t\t<ParticleOffsets/>\n\t\t\t<ExceptionOffsets/>")
            out.append("\t\t\t<Particles>")
            for i in range(f.particleCount):
                out.append(f'\t\t\t\t<Particle eps="{_n(f.epsilon[i])}" q="{_n(f.charge[i])}" '
                           f'sig="{_n(f.sigma[i])}"/>')
            out.append("\t\t\t</Particles>")
            out.append("\t\t\t<Exceptions>")
            for i in range(f.exceptionCount):
                out.append(f'\t\t\t\t<Exception eps="{_n(f.exceptionEpsilon[i])}" '
                           f'p1="{int(f.exceptionP1[i])}" p2="{int(f.exceptionP2[i])}" '
                           f'q="{_n(f.exceptionChargeProd[i])}" sig="{_n(f.exceptionSigma[i])}"/>')
            out.append("\t\t\t</Exceptions>")
        elif f.type == "CMMotionRemover":
            out.append(f'\t\t<Force frequency="{f.frequency}" {common} type="CMMotionRemover" version="1"/>')
            continue
        elif f.type == "MonteCarloBarostat":
            out.append(f'\t\t<Force frequency="{f.frequency}" {common} pressure="{_n(f.defaultPressure)}" '
                       f'randomSeed="{f.randomSeed}" temperature="{_n(f.defaultTemperature)}" '
                       f'type="MonteCarloBarostat" version="2"/>')
            continue
        elif f.type == "PositionRestraint":
            out.append(f'\t\t<Force energy="{f.energyExpression}" {common} '
                       f'type="CustomExternalForce" version="1">')
            out.append("\t\t\t<PerParticleParameters>")
            for pname in ("k", "x0", "y0", "z0"):
                out.append(f'\t\t\t\t<Parameter name="{pname}"/>')
            out.append("\t\t\t</PerParticleParameters>")
            out.append("\t\t\t<GlobalParameters/>")
            out.append("\t\t\t<Particles>")
            for i in range(f.count):
                out.append(f'\t\t\t\t<Particle index="{int(f.particle[i])}" param1="{_n(f.k[i])}" '
                           f'param2="{_n(f.x0[i])}" param3="{_n(f.y0[i])}" param4="{_n(f.z0[i])}"/>')
            out.append("\t\t\t</Particles>")
        out.append("\t\t</Force>")
    out.append("\t</Forces>")
    out.append("</System>")
    return "\n".join(out) + "\n"


# ---- box geometry used by the host (src/reference/triclinic.ts) ----------

def _reduce_by(ratio: float) -> float:
    nearest = math.floor(ratio + 0.5)
    if abs(ratio - math.trunc(ratio)) == 0.5 and nearest % 2 != 0:
        return nearest - math.copysign(1.0, ratio)
    return nearest


def reduceBox(box: Sequence[Sequence[float]]) -> BoxVectors:
    """OpenMM's reduced form, ties to even (``triclinic.ts:93-125``)."""
    v = [list(map(float, r)) for r in box]
    for target, source, num, den in ((2, 1, (2, 1), (1, 1)), (2, 0, (2, 0), (0, 0)),
                                     (1, 0, (1, 0), (0, 0))):
        t = _reduce_by(v[num[0]][num[1]] / v[den[0]][den[1]])
        v[target] = [v[target][k] - t * v[source][k] for k in range(3)]
    return (tuple(v[0]), tuple(v[1]), tuple(v[2]))


def isReducedForm(box: Sequence[Sequence[float]], tolerance: float = 1e-9) -> bool:
    a, b, c = box
    if abs(a[1]) > tolerance or abs(a[2]) > tolerance or abs(b[2]) > tolerance:
        return False
    if a[0] <= 0 or b[1] <= 0 or c[2] <= 0:
        return False
    return (abs(b[0]) <= 0.5 * a[0] + tolerance and abs(c[0]) <= 0.5 * a[0] + tolerance
            and abs(c[1]) <= 0.5 * b[1] + tolerance)


def boxVolume(box: Sequence[Sequence[float]]) -> float:
    return abs(float(np.linalg.det(np.asarray(box, dtype=np.float64))))


def perpendicularWidths(box: Sequence[Sequence[float]]) -> Vec3:
    a, b, c = (np.asarray(v, dtype=np.float64) for v in box)
    vol = boxVolume(box)
    return (vol / np.linalg.norm(np.cross(b, c)), vol / np.linalg.norm(np.cross(c, a)),
            vol / np.linalg.norm(np.cross(a, b)))


def boxVectorsForShape(shape: str, width: float) -> BoxVectors:
    """OpenMM's cube / rhombic dodecahedron / truncated octahedron cells
    (``triclinic.ts:359-392``)."""
    if shape == "cube":
        return ((width, 0.0, 0.0), (0.0, width, 0.0), (0.0, 0.0, width))
    if shape == "dodecahedron":
        return ((width, 0.0, 0.0), (0.0, width, 0.0),
                (width / 2, width / 2, width * math.sqrt(2) / 2))
    if shape == "octahedron":
        return ((width, 0.0, 0.0), (width / 3, width * 2 * math.sqrt(2) / 3, 0.0),
                (-width / 3, width * math.sqrt(2) / 3, width * math.sqrt(6) / 3))
    raise ValueError(f"unknown box shape {shape}")
