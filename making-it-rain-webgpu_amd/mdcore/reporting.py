"""The run log and the final-frame PDB of a stage.

* The log is OpenMM's ``StateDataReporter`` CSV with the notebook's flags
  (``src/analysis/log.ts:24-74``): a ``#``-prefixed header of quoted column
  names in OpenMM's fixed order, then one comma-separated row per report,
  ``--`` for a speed not yet known.
* ``rewritePdb`` substitutes positions (nm in, angstrom out, columns 31-54)
  into an uploaded topology PDB and replaces its ``CRYST1`` with the cell the
  stage ended at, prepending a ``REMARK 1`` line
  (``src/analysis/pdb.ts:141-188``).
"""
from __future__ import annotations

import math
from typing import Optional

import numpy as np

from .dcd import unitCellOf

LOG_COLUMNS = ("Step", "Potential Energy (kJ/mole)", "Kinetic Energy (kJ/mole)", "Total Energy (kJ/mole)",
               "Temperature (K)", "Box Volume (nm^3)", "Speed (ns/day)")


def _num(x) -> str:
    """Shortest decimal that reads back exactly, integers without a point
    (JavaScript's number formatting, which the upstream log uses)."""
    if isinstance(x, (int, np.integer)):
        return str(int(x))
    x = float(x)
    if math.isfinite(x) and x == int(x) and abs(x) < 1e21:
        return str(int(x))
    return repr(x)


def logHeader() -> str:
    return "#" + ",".join(f'"{c}"' for c in LOG_COLUMNS) + "\n"


def logRow(step: int, potentialEnergy: float, kineticEnergy: float, totalEnergy: float, temperature: float,
           volume: float, speed: Optional[float]) -> str:
    return ",".join([_num(step), _num(potentialEnergy), _num(kineticEnergy), _num(totalEnergy), _num(temperature),
                     _num(volume), "--" if speed is None else _num(speed)]) + "\n"


class PdbAtomCountMismatch(ValueError):
    def __init__(self, in_file: int, atoms: int):
        super().__init__(f"the PDB has {in_file} atom records but there are {atoms} positions: the two have to "
                         "describe the same atoms in the same order")


def _is_atom(line: str) -> bool:
    return line.startswith("ATOM") or line.startswith("HETATM")


def _fixed(value: float, width: int, decimals: int) -> str:
    text = f"{value:.{decimals}f}"
    if len(text) > width:
        raise ValueError(f"{text} does not fit the {width}-column PDB field it belongs in")
    return text.rjust(width)


def rewritePdb(pdb: str, positions, box=None, remark: Optional[str] = None) -> str:
    x = np.asarray(positions, dtype=np.float64).reshape(-1)
    atoms = x.shape[0] // 3
    in_file = sum(1 for line in pdb.split("\n") if _is_atom(line))
    if in_file != atoms:
        raise PdbAtomCountMismatch(in_file, atoms)
    out = []
    if remark is not None:
        out.append(f"REMARK   1 {remark}".ljust(80)[:80])
    k = 0
    for line in pdb.split("\n"):
        if _is_atom(line):
            padded = line.ljust(54)
            out.append(padded[:30] + _fixed(x[3 * k] * 10.0, 8, 3) + _fixed(x[3 * k + 1] * 10.0, 8, 3)
                       + _fixed(x[3 * k + 2] * 10.0, 8, 3) + padded[54:])
            k += 1
            continue
        if line.startswith("CRYST1") and box is not None:
            cell = unitCellOf(box)
            a, b, c = (v * 10.0 for v in cell.lengths)
            al, be, ga = (math.degrees(math.acos(max(-1.0, min(1.0, v)))) for v in cell.cosines)
            out.append("CRYST1" + _fixed(a, 9, 3) + _fixed(b, 9, 3) + _fixed(c, 9, 3) + _fixed(al, 7, 2)
                       + _fixed(be, 7, 2) + _fixed(ga, 7, 2) + " P 1           1 ")
            continue
        out.append(line)
    return "\n".join(out)
