"""DCD trajectories in the CHARMM-24 little-endian dialect OpenMM writes.

Mirrors ``src/analysis/dcd.ts``: ``DcdWriter`` (header with patched frame
count and last step, one record set per frame), ``readDcd`` (strict: every
Fortran record marker is checked) and ``unitCellOf``.  The format's fixed
points: coordinates and edge lengths in angstrom, the timestep in AKMA units
(``dt / 0.04888821`` ps), the unit-cell record ordered
``[a, cos gamma, b, cos beta, cos alpha, c]``, a 276-byte header
(``dcd.ts:112``), ``NSET`` at byte 8 and ``NSTEP`` at byte 20.

Layout written here, per the CHARMM record structure (every record is
``int32 length, payload, int32 length``):

* header record (84 bytes): ``"CORD"``, icntrl[20] with icntrl[0] NSET,
  [1] ISTART (step of the first frame), [2] NSAVC (interval), [3] NSTEP (step
  of the last frame), [9] DELTA as float32, [10] unit cell flag, [19] 24;
* title record (164 bytes): NTITLE = 2 and two 80-byte NUL-padded lines;
* atom count record (4 bytes);
* per frame: the 48-byte unit cell record (periodic files only), then one
  record of float32 per axis.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

AKMA_PICOSECONDS = 0.04888821  # one AKMA time unit in ps (dcd.ts:91)
TO_ANGSTROM = 10.0
DCD_HEADER_BYTES = 276         # dcd.ts:112
NSET_OFFSET = 8
NSTEP_OFFSET = 20


@dataclass(frozen=True)
class UnitCell:
    """Edge lengths (nm) and ``[cos alpha, cos beta, cos gamma]``."""

    lengths: tuple
    cosines: tuple


def unitCellOf(box) -> UnitCell:
    """The cell OpenMM writes for three box vectors (``dcd.ts:71-73``): edge
    lengths and the cosines as dot-product ratios (alpha between b and c),
    with no trip through acos, so a dodecahedron's 0.5 stays exactly 0.5."""
    a, b, c = (np.asarray(v, dtype=np.float64) for v in box)
    la, lb, lc = (float(np.sqrt(np.dot(v, v))) for v in (a, b, c))
    return UnitCell((la, lb, lc), (float(np.dot(b, c)) / (lb * lc), float(np.dot(c, a)) / (lc * la),
                                   float(np.dot(a, b)) / (la * lb)))


class DcdWriter:
    """``DcdWriter`` (``dcd.ts:114-297``)."""

    def __init__(self, particleCount: int, timeStep: float, firstStep: int = 0, interval: int = 1,
                 periodic: bool = True, titles: Optional[Sequence[str]] = None):
        if int(particleCount) != particleCount or particleCount <= 0:
            raise ValueError(f"a DCD needs at least one particle, got {particleCount}")
        self.particleCount = int(particleCount)
        self.timeStep = float(timeStep)
        self.firstStep = int(firstStep)
        self.interval = int(interval)
        self.periodic = bool(periodic)
        self.titles = list(titles) if titles is not None else ["Created by making-it-rain-webgpu"]
        self.frames = 0

    @property
    def frameCount(self) -> int:
        return self.frames

    @property
    def frameBytes(self) -> int:
        return 3 * (8 + 4 * self.particleCount) + (56 if self.periodic else 0)

    def header(self) -> bytes:
        icntrl = [0] * 20
        icntrl[1] = self.firstStep
        icntrl[2] = self.interval
        icntrl[10] = 1 if self.periodic else 0
        icntrl[19] = 24  # CHARMM version: selects this dialect
        body = b"CORD" + struct.pack("<9i", *icntrl[:9]) + struct.pack("<f", self.timeStep / AKMA_PICOSECONDS) \
            + struct.pack("<10i", *icntrl[10:])
        out = struct.pack("<i", 84) + body + struct.pack("<i", 84)
        lines = b""
        for k in range(2):
            text = self.titles[k] if k < len(self.titles) else ""
            raw = bytes(ord(ch) if ord(ch) < 128 else 0x3F for ch in text[:80])
            lines += raw.ljust(80, b"\0")
        out += struct.pack("<ii", 164, 2) + lines + struct.pack("<i", 164)
        out += struct.pack("<iii", 4, self.particleCount, 4)
        assert len(out) == DCD_HEADER_BYTES
        return out

    def frame(self, positions, box=None) -> bytes:
        x = np.asarray(positions, dtype=np.float64).reshape(-1)
        n = self.particleCount
        if x.shape[0] != 3 * n:
            raise ValueError(f"expected {3 * n} coordinates for {n} particles, got {x.shape[0]}")
        if self.periodic and box is None:
            raise ValueError("a periodic DCD needs a box for every frame; the header says each frame carries one")
        if not np.all(np.isfinite(x)):
            bad = int(np.argmax(~np.isfinite(x)))
            raise ValueError(f"coordinate {bad} is {x[bad]}: a non-finite position (NaN or infinity) is not written "
                             "to a trajectory")
        out = b""
        if self.periodic:
            cell = unitCellOf(box)
            values = (cell.lengths[0] * TO_ANGSTROM, cell.cosines[2], cell.lengths[1] * TO_ANGSTROM,
                      cell.cosines[1], cell.cosines[0], cell.lengths[2] * TO_ANGSTROM)
            out += struct.pack("<i6di", 48, *values, 48)
        marker = struct.pack("<i", 4 * n)
        xyz = (x.reshape(n, 3) * TO_ANGSTROM).astype("<f4")
        for axis in range(3):
            out += marker + np.ascontiguousarray(xyz[:, axis]).tobytes() + marker
        self.frames += 1
        return out

    def headerPatches(self):
        """(offset, bytes) of NSET and NSTEP once the frame count is known."""
        last = self.firstStep + max(0, self.frames - 1) * self.interval
        return [(NSET_OFFSET, struct.pack("<i", self.frames)), (NSTEP_OFFSET, struct.pack("<i", last))]

    def patch(self, data: bytearray) -> None:
        for offset, word in self.headerPatches():
            data[offset:offset + 4] = word


class DcdFormatError(ValueError):
    """A file this reader would misread (``dcd.ts:330-341``)."""

    def __init__(self, detail: str):
        super().__init__(f"this does not look like a DCD this reader can handle: {detail}")


@dataclass
class DcdTrajectory:
    particleCount: int
    frameCount: int
    firstStep: int
    interval: int
    timeStep: float
    periodic: bool
    positions: np.ndarray  # (frames, n, 3) nm, float32
    boxes: np.ndarray      # (frames, 3) edge lengths nm
    cosines: np.ndarray    # (frames, 3) cos alpha, beta, gamma


def readDcd(data: bytes) -> DcdTrajectory:
    """``readDcd`` (``dcd.ts:361-492``)."""
    buf = memoryview(bytes(data))
    if len(buf) < DCD_HEADER_BYTES:
        raise DcdFormatError(f"only {len(buf)} bytes, less than a header")

    def i32(at):
        return struct.unpack_from("<i", buf, at)[0]

    first = i32(0)
    if first != 84:
        if struct.unpack_from(">i", buf, 0)[0] == 84:
            raise DcdFormatError("the leading record marker is big-endian; only the little-endian dialect OpenMM "
                                 "writes is supported")
        raise DcdFormatError(f"the leading record marker is {first}, not 84")
    magic = bytes(buf[4:8]).decode("latin-1")
    if magic != "CORD":
        raise DcdFormatError(f'the magic is "{magic}", not "CORD"')
    count_hdr, first_step, interval = i32(8), i32(12), i32(16)
    delta = struct.unpack_from("<f", buf, 44)[0]
    periodic = i32(48) != 0
    if i32(84) == 0:
        raise DcdFormatError("the CHARMM version word is zero, which marks an X-PLOR style file")
    if i32(88) != 84:
        raise DcdFormatError("the header's closing record marker is not 84")
    title_bytes = i32(92)
    at = 96 + title_bytes
    if i32(at) != title_bytes:
        raise DcdFormatError("the title record's two markers disagree")
    at += 4
    if i32(at) != 4:
        raise DcdFormatError("the atom-count record is not four bytes long")
    n = i32(at + 4)
    if n <= 0:
        raise DcdFormatError(f"{n} particles")
    at += 12
    per_frame = 3 * (4 * n + 8) + (56 if periodic else 0)
    frames = (len(buf) - at) // per_frame
    if frames == 0:
        raise DcdFormatError(f"no complete frame: {len(buf) - at} bytes left but a frame needs {per_frame}")
    # a run killed mid-write leaves NSET claiming more frames than were flushed
    count = min(count_hdr, frames) if count_hdr > 0 else frames
    pos = np.zeros((count, n, 3), dtype=np.float32)
    boxes = np.zeros((count, 3)) if periodic else np.zeros((0, 3))
    cos = np.zeros((count, 3)) if periodic else np.zeros((0, 3))
    for f in range(count):
        if periodic:
            if i32(at) != 48:
                raise DcdFormatError(f"frame {f}'s unit-cell record is not 48 bytes")
            a, cg, b, cb, ca, c = struct.unpack_from("<6d", buf, at + 4)
            if i32(at + 52) != 48:
                raise DcdFormatError(f"frame {f}'s unit-cell record markers disagree")
            boxes[f] = (a / TO_ANGSTROM, b / TO_ANGSTROM, c / TO_ANGSTROM)
            cos[f] = (ca, cb, cg)
            at += 56
        for axis in range(3):
            if i32(at) != 4 * n:
                raise DcdFormatError(f"frame {f} axis {axis}: the record marker is {i32(at)}, not {4 * n}")
            pos[f, :, axis] = np.frombuffer(buf, dtype="<f4", count=n, offset=at + 4) / np.float32(TO_ANGSTROM)
            if i32(at + 4 + 4 * n) != 4 * n:
                raise DcdFormatError(f"frame {f} axis {axis}: markers disagree")
            at += 8 + 4 * n
    return DcdTrajectory(n, count, first_step, interval, float(delta) * AKMA_PICOSECONDS, periodic, pos, boxes, cos)
