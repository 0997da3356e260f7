"""The engine's other input: an OpenMM ``State`` (``src/engine/state.ts``).

Positions, box vectors and optionally velocities and forces, as float64 so no
precision is lost before the engine converts positions to fixed point
(``state.ts:1-13``).  ``serialiseState`` writes the ``.rst`` checkpoint format
(``state.ts:203-235``).
"""
from __future__ import annotations

import xml.etree.ElementTree as ET
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .system import BoxVectors, XmlError, _box, _n


@dataclass
class OpenMMState:
    """``OpenMMState`` (``state.ts:18-39``)."""

    boxVectors: BoxVectors
    positions: np.ndarray
    velocities: Optional[np.ndarray] = None
    forces: Optional[np.ndarray] = None
    stepCount: int = 0
    time: float = 0.0
    potentialEnergy: Optional[float] = None
    kineticEnergy: Optional[float] = None
    openmmVersion: str = "8.4"

    @property
    def particleCount(self) -> int:
        return int(self.positions.shape[0] // 3)


def _vectors(root: ET.Element, container: str, element: str) -> Optional[np.ndarray]:
    c = root.find(container)
    if c is None:
        return None
    values = []
    for v in c.findall(element):
        values.extend((float(v.get("x")), float(v.get("y")), float(v.get("z"))))
    return np.asarray(values, dtype=np.float64)


def parseState(xml: str) -> OpenMMState:
    """Parse State XML (``state.ts:60-152``)."""
    try:
        root = ET.fromstring(xml)
    except ET.ParseError as error:
        raise XmlError(str(error)) from error
    if root.tag != "State":
        raise XmlError("no <State> element")
    positions = _vectors(root, "Positions", "Position")
    if positions is None:
        raise XmlError("no <Positions>")
    energies = root.find("Energies")
    pe = ke = None
    if energies is not None:
        if energies.get("PotentialEnergy") is not None:
            pe = float(energies.get("PotentialEnergy"))
        if energies.get("KineticEnergy") is not None:
            ke = float(energies.get("KineticEnergy"))
    state = OpenMMState(
        boxVectors=_box(root.find("PeriodicBoxVectors")),
        positions=positions,
        velocities=_vectors(root, "Velocities", "Velocity"),
        forces=_vectors(root, "Forces", "Force"),
        stepCount=int(float(root.get("stepCount", "0"))),
        time=float(root.get("time", "0")),
        potentialEnergy=pe,
        kineticEnergy=ke,
        openmmVersion=root.get("openmmVersion", ""),
    )
    for name in ("velocities", "forces"):
        arr = getattr(state, name)
        if arr is not None and arr.shape[0] != positions.shape[0]:
            raise XmlError(f"{name} has {arr.shape[0] // 3} entries but there are "
                           f"{positions.shape[0] // 3} positions")
    return state


def _vector_list(container: str, element: str, values: np.ndarray) -> str:
    out = [f"\t<{container}>"]
    for i in range(0, values.shape[0], 3):
        out.append(f'\t\t<{element} x="{_n(values[i])}" y="{_n(values[i + 1])}" '
                   f'z="{_n(values[i + 2])}"/>')
    out.append(f"\t</{container}>")
    return "\n".join(out) + "\n"


def serialiseState(state: OpenMMState, openmmVersion: Optional[str] = None,
                   includeForces: bool = False) -> str:
    """Write State XML, the ``.rst`` checkpoint format (``state.ts:203-235``)."""
    version = openmmVersion or state.openmmVersion or "8.4"
    xml = '<?xml version="1.0" ?>\n'
    xml += (f'<State openmmVersion="{version}" stepCount="{state.stepCount}" '
            f'time="{_n(state.time)}" type="State" version="1">\n')
    xml += "\t<PeriodicBoxVectors>\n"
    for label, v in zip("ABC", state.boxVectors):
        xml += f'\t\t<{label} x="{_n(v[0])}" y="{_n(v[1])}" z="{_n(v[2])}"/>\n'
    xml += "\t</PeriodicBoxVectors>\n"
    xml += _vector_list("Positions", "Position", state.positions)
    if state.velocities is not None:
        xml += _vector_list("Velocities", "Velocity", state.velocities)
    if includeForces and state.forces is not None:
        xml += _vector_list("Forces", "Force", state.forces)
    xml += "</State>\n"
    return xml
