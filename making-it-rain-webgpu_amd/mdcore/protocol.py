"""The run: minimise, equilibrate, produce by strides (``src/engine/protocol.ts``).

The notebook's protocol as a state machine over any ``ProtocolEngine``
(``protocol.ts:32-62``; ``HipEngine`` implements it on the GPU) and a
``RunOutputs`` sink (append / patch / write / read / exists):

* ``runStage`` (``protocol.ts:216-355``) steps in chunks that end where the
  next trajectory frame or log row is due, on the **absolute** step counter
  (OpenMM's ``describeNextReport``), so nothing is reported at step zero and a
  stride that starts mid-interval keeps the cadence.  It writes a DCD (first
  frame's step as ``ISTART``, ``NSET``/``NSTEP`` patched at the end), the
  ``StateDataReporter`` log (rewritten whole at each row), the ``.rst``
  checkpoint (``serialiseState`` of the whole cell, positions and velocities)
  and, given a topology PDB, the final frame.
* ``runProtocol`` (``protocol.ts:393-517``): equilibration (restrained
  minimisation, then velocities, then dynamics) unless its checkpoint exists,
  then production strides named ``<job>_prod_<k>``, each from the previous
  checkpoint on a fresh engine, a stride whose checkpoint exists skipped;
  ``run.json`` records every setting, the seed and every stage, skipped ones
  included.

One deliberate difference: a stage that runs starts its trajectory afresh
(``RunOutputs.write_bytes``) instead of appending its header to whatever an
interrupted earlier attempt at the same stage left behind, which upstream's
append would turn into a file with two headers.
"""
from __future__ import annotations

import dataclasses
import datetime
import json
import os
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

from .dcd import DcdWriter
from .reporting import logHeader, logRow, rewritePdb
from .state import OpenMMState, parseState, serialiseState


# ---- where artefacts go ------------------------------------------------------

class MemoryOutputs:
    """An in-memory ``RunOutputs`` (what the upstream tests use)."""

    def __init__(self):
        self.files: Dict[str, bytearray] = {}

    def append(self, name: str, data: bytes) -> None:
        self.files.setdefault(name, bytearray()).extend(data)

    def write_bytes(self, name: str, data: bytes) -> None:
        self.files[name] = bytearray(data)

    def patch(self, name: str, offset: int, data: bytes) -> None:
        if name not in self.files:
            raise FileNotFoundError(f"cannot patch missing {name}")
        self.files[name][offset:offset + len(data)] = data

    def write(self, name: str, text: str) -> None:
        self.files[name] = bytearray(text.encode())

    def read(self, name: str) -> Optional[str]:
        data = self.files.get(name)
        return None if data is None else bytes(data).decode()

    def exists(self, name: str) -> bool:
        return name in self.files


class DirectoryOutputs:
    """A ``RunOutputs`` over a directory of files.  Text files are written to a
    temporary name and renamed over the old one, so a run killed mid-write
    leaves the previous complete file."""

    def __init__(self, directory: str):
        self.directory = directory
        os.makedirs(directory, exist_ok=True)

    def _path(self, name: str) -> str:
        return os.path.join(self.directory, name)

    def append(self, name: str, data: bytes) -> None:
        with open(self._path(name), "ab") as f:
            f.write(data)

    def write_bytes(self, name: str, data: bytes) -> None:
        with open(self._path(name), "wb") as f:
            f.write(data)

    def patch(self, name: str, offset: int, data: bytes) -> None:
        with open(self._path(name), "r+b") as f:
            f.seek(offset)
            f.write(data)

    def write(self, name: str, text: str) -> None:
        tmp = self._path(name) + ".partial"
        with open(tmp, "w") as f:
            f.write(text)
        os.replace(tmp, self._path(name))

    def read(self, name: str) -> Optional[str]:
        try:
            with open(self._path(name)) as f:
                return f.read()
        except FileNotFoundError:
            return None

    def exists(self, name: str) -> bool:
        return os.path.exists(self._path(name))


# ---- settings ----------------------------------------------------------------

@dataclass
class StageSettings:
    """``StageSettings`` (``protocol.ts:82-95``): ns, ps, K, bar, ps, ps."""

    nanoseconds: float
    timeStep: float
    temperature: float
    pressure: float
    trajectoryInterval: float
    logInterval: float


@dataclass
class EquilibrationSettings(StageSettings):
    minimisationSteps: int = 0
    restraintForceConstant: float = 0.0
    minimisationTolerance: float = 10.0


@dataclass
class ProductionSettings(StageSettings):
    strides: int = 1


@dataclass
class ProtocolSettings:
    jobName: str
    equilibration: EquilibrationSettings
    production: ProductionSettings
    seed: int = 0


@dataclass
class StageResult:
    name: str
    steps: int
    frames: int
    logRows: int
    elapsed: float
    nsPerDay: float
    skipped: bool


def _skipped(name: str) -> StageResult:
    return StageResult(name, 0, 0, 0, 0.0, 0.0, True)


def stepsForInterval(picoseconds: float, timeStep: float, what: str) -> int:
    """Steps in an interval; refuses a non-integer rather than rounding, which
    would drift every frame time (``protocol.ts:110-137``)."""
    steps = picoseconds / timeStep
    rounded = round(steps)
    if abs(steps - rounded) > 1e-9 or rounded < 1:
        raise ValueError(f"{what} of {picoseconds} ps is {steps} steps at a {timeStep} ps timestep, which is not a "
                         "whole number. Choose an interval that divides by the timestep: otherwise the frame times "
                         "drift from the step numbers by a growing amount, and nothing reports it.")
    return int(rounded)


def stepsForStage(nanoseconds: float, timeStep: float) -> int:
    return stepsForInterval(nanoseconds * 1000.0, timeStep, "a stage length")


def artefactNames(base: str) -> Dict[str, str]:
    return {"trajectory": f"{base}.dcd", "log": f"{base}.log", "checkpoint": f"{base}.rst",
            "finalFrame": f"{base}.pdb"}


def stageNames(settings: ProtocolSettings) -> Dict[str, object]:
    return {"equilibration": f"{settings.jobName}_equil",
            "strides": [f"{settings.jobName}_prod_{k + 1}" for k in range(settings.production.strides)]}


def _sample(engine):
    s = engine.sample()
    get = (lambda k: s[k]) if isinstance(s, dict) else (lambda k: getattr(s, k))
    return get("potentialEnergy"), get("kineticEnergy"), get("totalEnergy"), get("temperature")


def runStage(engine, outputs, base: str, settings: StageSettings, resume: bool = False,
             topologyPdb: Optional[str] = None, onReport: Optional[Callable[[dict], None]] = None) -> StageResult:
    """One stage: trajectory, log, checkpoint (``protocol.ts:216-355``)."""
    names = artefactNames(base)
    if resume and outputs.exists(names["checkpoint"]):
        return _skipped(base)
    total = stepsForStage(settings.nanoseconds, settings.timeStep)
    frame_every = stepsForInterval(settings.trajectoryInterval, settings.timeStep, "a trajectory interval")
    log_every = stepsForInterval(settings.logInterval, settings.timeStep, "a log interval")
    writer = DcdWriter(engine.particleCount, settings.timeStep, firstStep=engine.steps + frame_every,
                       interval=frame_every, periodic=True,
                       titles=[f"Created by making-it-rain-webgpu: {base}", ""])
    outputs.write_bytes(names["trajectory"], writer.header())
    log = logHeader()
    outputs.write(names["log"], log)
    began = time.perf_counter()
    final = engine.steps + total
    frames = rows = 0
    while engine.steps < final:
        until_frame = frame_every - engine.steps % frame_every
        until_log = log_every - engine.steps % log_every
        engine.run(min(until_frame, until_log, final - engine.steps))
        elapsed = time.perf_counter() - began
        if engine.steps % frame_every == 0:
            outputs.append(names["trajectory"], writer.frame(engine.positions(), engine.boxVectors()))
            frames += 1
        if engine.steps % log_every == 0:
            pe, ke, te, temp = _sample(engine)
            bl = engine.boxLengths()
            volume = bl[0] * bl[1] * bl[2]
            done = engine.steps - (final - total)
            ns_day = (done * settings.timeStep / 1000.0) / elapsed * 86400.0 if elapsed > 0 else None
            log += logRow(engine.steps, pe, ke, te, temp, volume, ns_day)
            rows += 1
            if onReport is not None:
                onReport({"step": engine.steps, "totalSteps": final, "potentialEnergy": pe, "temperature": temp,
                          "nsPerDay": ns_day or 0.0})
            outputs.write(names["log"], log)
    for offset, word in writer.headerPatches():
        outputs.patch(names["trajectory"], offset, word)
    positions = engine.positions()
    checkpoint = OpenMMState(boxVectors=engine.boxVectors(), positions=positions, velocities=engine.velocities(),
                             stepCount=engine.steps, time=engine.steps * settings.timeStep, openmmVersion="8.4")
    outputs.write(names["checkpoint"], serialiseState(checkpoint))
    if topologyPdb is not None:
        outputs.write(names["finalFrame"], rewritePdb(topologyPdb, positions, box=engine.boxVectors(),
                                                      remark=f"FINAL FRAME OF {base.upper()}, STEP {engine.steps}"))
    elapsed = time.perf_counter() - began
    ns = total * settings.timeStep / 1000.0
    return StageResult(base, total, frames, rows, elapsed, ns / elapsed * 86400.0 if elapsed > 0 else 0.0, False)


@dataclass
class ProtocolHooks:
    """``ProtocolHooks`` (``protocol.ts:371-391``): the caller builds the engine
    each stage needs (restrained for equilibration, a fresh unrestrained
    system for production) and draws the initial velocities."""

    engineFor: Callable[[str, Optional[OpenMMState]], object]
    initialVelocities: Callable[[object], object]
    topologyPdb: Optional[str] = None
    onStage: Optional[Callable[[StageResult], None]] = None
    onReport: Optional[Callable[[dict], None]] = None


def _now() -> str:
    return datetime.datetime.now(datetime.timezone.utc).isoformat()


def runProtocol(system, outputs, settings: ProtocolSettings, hooks: ProtocolHooks) -> dict:
    """The whole protocol; returns (and writes as ``run.json``) the run record."""
    names = stageNames(settings)
    stages: List[StageResult] = []
    started = _now()
    minimised = None
    eq_names = artefactNames(names["equilibration"])
    if outputs.exists(eq_names["checkpoint"]):
        stages.append(_skipped(names["equilibration"]))
    else:
        engine = hooks.engineFor("equilibration", None)
        try:
            minimised = engine.minimise(settings.equilibration.minimisationSteps,
                                        settings.equilibration.minimisationTolerance)
            engine.setVelocities(hooks.initialVelocities(engine))
            result = runStage(engine, outputs, names["equilibration"], settings.equilibration,
                              topologyPdb=hooks.topologyPdb, onReport=hooks.onReport)
        finally:
            _release(engine)
        stages.append(result)
        if hooks.onStage:
            hooks.onStage(result)
    carried = None
    for index, base in enumerate(names["strides"]):
        if outputs.exists(artefactNames(base)["checkpoint"]):
            skipped = _skipped(base)
            stages.append(skipped)
            if hooks.onStage:
                hooks.onStage(skipped)
            carried = None
            continue
        if carried is None:
            previous = eq_names["checkpoint"] if index == 0 else artefactNames(names["strides"][index - 1])[
                "checkpoint"]
            text = outputs.read(previous)
            if text is None:
                raise RuntimeError(f"{base} needs {previous} to start from, and it is not there. A production stride "
                                   "continues the previous one's state; running them out of order is not supported.")
            carried = parseState(text)
        engine = hooks.engineFor("production", carried)
        try:
            result = runStage(engine, outputs, base, settings.production, topologyPdb=hooks.topologyPdb,
                              onReport=hooks.onReport)
        finally:
            _release(engine)
        stages.append(result)
        if hooks.onStage:
            hooks.onStage(result)
        carried = None
    artefacts = list(eq_names.values())
    for base in names["strides"]:
        artefacts += list(artefactNames(base).values())
    artefacts.append("run.json")
    record = {
        "jobName": settings.jobName,
        "seed": settings.seed,
        "settings": dataclasses.asdict(settings),
        "particleCount": system.particleCount,
        "stages": [dataclasses.asdict(s) for s in stages],
        "minimisedEnergy": minimised,
        "artefacts": artefacts,
        "startedAt": started,
        "finishedAt": _now(),
    }
    outputs.write("run.json", json.dumps(record, indent=1) + "\n")
    return record


def _release(engine) -> None:
    destroy = getattr(engine, "destroy", None)
    if callable(destroy):
        destroy()
