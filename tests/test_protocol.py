"""The protocol's orchestration against a stub engine (CPU).

Mirrors the upstream tests of ``src/engine/protocol.ts``
(``test/protocol.test.ts:147-343``): when reports land on the absolute step
counter, which header words the trajectory gets, what the checkpoint holds,
which stages a resumed run skips and what ``run.json`` records.  No dynamics:
a stub makes the orchestration deterministic.
"""
import json
import struct

import numpy as np
import pytest

from mdcore.protocol import (DirectoryOutputs, EquilibrationSettings, MemoryOutputs, ProductionSettings,
                             ProtocolHooks, ProtocolSettings, StageSettings, artefactNames, runProtocol, runStage,
                             stageNames, stepsForInterval, stepsForStage)
from mdcore.state import parseState


class StubEngine:
    """Counts steps and reports plausible numbers (``protocol.test.ts:61-124``)."""

    def __init__(self, n, start=0):
        self.particleCount = n
        self.steps = start
        self.calls = []
        self._pos = np.array([(i % 7) * 0.3 for i in range(3 * n)], dtype=np.float64)
        self._vel = np.zeros(3 * n)
        self._box = (3.0, 3.0, 3.0)

    def positions(self):
        return self._pos

    def velocities(self):
        return self._vel

    def boxLengths(self):
        return self._box

    def boxVectors(self):
        b = self._box
        return ((b[0], 0.0, 0.0), (0.0, b[1], 0.0), (0.0, 0.0, b[2]))

    def setPositions(self, x):
        self.calls.append("setPositions")

    def setVelocities(self, v):
        self.calls.append("setVelocities")

    def setBox(self, b):
        self._box = tuple(b)

    def run(self, count):
        self.calls.append(f"run:{count}")
        self.steps += count

    def sample(self):
        return {"potentialEnergy": -1000 - self.steps, "kineticEnergy": 500, "totalEnergy": -500 - self.steps,
                "temperature": 298}

    def minimise(self, max_iterations, tolerance):
        self.calls.append(f"minimise:{max_iterations}")
        return -12345


TOPOLOGY = "\n".join([
    "CRYST1   30.000   30.000   30.000  90.00  90.00  90.00 P 1           1 ",
    "ATOM      1  N   ASN A   1       1.000   2.000   3.000  1.00  0.00           N  ",
    "ATOM      2  CA  ASN A   1       1.500   2.000   3.000  1.00  0.00           C  ",
    "ATOM      3  C   ASN A   1       2.000   2.000   3.000  1.00  0.00           C  ",
    "ATOM      4  O   ASN A   1       2.500   2.000   3.000  1.00  0.00           O  ",
    "END",
])

# 10,000 steps at 2 fs, frames every 2,000 steps, rows every 1,000
STAGE = dict(nanoseconds=0.02, timeStep=0.002, temperature=298, pressure=1, trajectoryInterval=4, logInterval=2)


def stage():
    return StageSettings(**STAGE)


def test_interval_arithmetic():
    assert stepsForInterval(10, 0.002, "x") == 5000
    assert stepsForStage(1, 0.002) == 500_000
    with pytest.raises(ValueError, match="not a whole number"):
        stepsForInterval(10, 0.003, "a trajectory interval")
    with pytest.raises(ValueError, match="whole number"):
        stepsForInterval(0.001, 0.002, "x")


def test_reports_on_multiples_of_the_interval_and_not_at_step_zero():
    out = MemoryOutputs()
    res = runStage(StubEngine(4), out, "job_equil", stage())
    assert (res.steps, res.frames, res.logRows) == (10_000, 5, 10)
    rows = out.read("job_equil.log").strip().split("\n")
    assert rows[0] == ('#"Step","Potential Energy (kJ/mole)","Kinetic Energy (kJ/mole)","Total Energy (kJ/mole)",'
                       '"Temperature (K)","Box Volume (nm^3)","Speed (ns/day)"')
    assert rows[1].startswith("1000,") and rows[-1].startswith("10000,")
    assert rows[1].split(",")[1:6] == ["-2000", "500", "-1500", "298", "27"]


def test_cadence_follows_the_absolute_step_counter():
    out = MemoryOutputs()
    runStage(StubEngine(4, 10_000), out, "job_prod_2", stage())
    rows = out.read("job_prod_2.log").strip().split("\n")[1:]
    assert rows[0].startswith("11000,") and rows[-1].startswith("20000,")
    # a stage that starts mid-interval reports on the absolute multiples
    out2 = MemoryOutputs()
    e = StubEngine(4, 1_500)
    runStage(e, out2, "job_prod_x", stage())
    rows = out2.read("job_prod_x.log").strip().split("\n")[1:]
    assert rows[0].startswith("2000,") and rows[-1].startswith("11000,")
    assert e.calls[0] == "run:500"


def test_trajectory_header_knows_its_frames():
    out = MemoryOutputs()
    res = runStage(StubEngine(4), out, "job_equil", stage())
    dcd = bytes(out.files["job_equil.dcd"])
    nset, istart, nsavc, nstep = struct.unpack_from("<4i", dcd, 8)
    assert nset == res.frames == 5
    assert istart == 2_000      # the first frame's step, not the stage start
    assert nsavc == 2_000
    assert nstep == 10_000      # the last frame's step


def test_final_frame_pdb_only_with_a_topology():
    out = MemoryOutputs()
    runStage(StubEngine(4), out, "job_bare", stage())
    assert not out.exists("job_bare.pdb")
    runStage(StubEngine(4), out, "job_equil", stage(), topologyPdb=TOPOLOGY)
    pdb = out.read("job_equil.pdb")
    assert len([line for line in pdb.split("\n") if line.startswith("ATOM")]) == 4
    assert "FINAL FRAME OF JOB_EQUIL, STEP 10000" in pdb
    assert "CRYST1   30.000   30.000   30.000  90.00  90.00  90.00" in pdb
    # positions in angstrom in columns 31-54: atom 2 is (0.9, 1.2, 1.5) nm
    atom2 = [line for line in pdb.split("\n") if line.startswith("ATOM")][1]
    assert atom2[30:54] == "   9.000  12.000  15.000"


def test_checkpoint_is_a_state_openmm_reads():
    out = MemoryOutputs()
    runStage(StubEngine(4), out, "job_equil", stage())
    cp = parseState(out.read("job_equil.rst"))
    assert cp.stepCount == 10_000
    assert cp.time == pytest.approx(20.0, abs=1e-9)
    assert cp.particleCount == 4
    assert cp.velocities is not None


def test_resume_skips_a_stage_whose_checkpoint_exists():
    out = MemoryOutputs()
    out.write("job_prod_1.rst", "<State/>")
    e = StubEngine(4)
    res = runStage(e, out, "job_prod_1", stage(), resume=True)
    assert res.skipped and e.calls == []


def _settings():
    return ProtocolSettings(
        jobName="job",
        equilibration=EquilibrationSettings(**STAGE, minimisationSteps=10_000, restraintForceConstant=700),
        production=ProductionSettings(**STAGE, strides=2),
        seed=0x5EED)


class _System:
    particleCount = 4


def _hooks(engines):
    def engine_for(which, state):
        e = StubEngine(4, 0 if which == "equilibration" else 10_000)
        engines.append(e)
        return e
    return ProtocolHooks(engineFor=engine_for, initialVelocities=lambda e: np.zeros(3 * e.particleCount),
                         topologyPdb=TOPOLOGY)


def test_minimises_before_drawing_velocities_and_before_dynamics():
    engines = []
    runProtocol(_System(), MemoryOutputs(), _settings(), _hooks(engines))
    calls = engines[0].calls
    assert calls[0] == "minimise:10000"
    assert calls[1] == "setVelocities"
    assert calls[2].startswith("run:")


def test_runs_every_stride_and_names_them_as_the_notebook_does():
    out = MemoryOutputs()
    rec = runProtocol(_System(), out, _settings(), _hooks([]))
    assert stageNames(_settings())["strides"] == ["job_prod_1", "job_prod_2"]
    assert [s["name"] for s in rec["stages"]] == ["job_equil", "job_prod_1", "job_prod_2"]
    for base in ("job_equil", "job_prod_1", "job_prod_2"):
        for name in artefactNames(base).values():
            assert out.exists(name), name


def test_records_the_run_including_what_it_skipped():
    out = MemoryOutputs()
    runProtocol(_System(), out, _settings(), _hooks([]))
    rec = json.loads(out.read("run.json"))
    assert rec["seed"] == 0x5EED
    assert rec["minimisedEnergy"] == -12345
    assert rec["settings"]["equilibration"]["restraintForceConstant"] == 700
    again = runProtocol(_System(), out, _settings(), _hooks([]))
    assert all(s["skipped"] for s in again["stages"])
    assert len(again["stages"]) == 3
    assert again["minimisedEnergy"] is None


def test_stride_starts_from_the_previous_checkpoint():
    seen = []

    def engine_for(which, state):
        seen.append(None if state is None else f"state:{state.stepCount}")
        return StubEngine(4, 0 if which == "equilibration" else 10_000)

    runProtocol(_System(), MemoryOutputs(), _settings(),
                ProtocolHooks(engineFor=engine_for, initialVelocities=lambda e: np.zeros(12)))
    assert seen == [None, "state:10000", "state:20000"]


def test_refuses_a_stride_whose_predecessor_left_no_checkpoint():
    out = MemoryOutputs()
    out.write("job_equil.rst", "<State/>")
    real_read = out.read
    out.read = lambda name: None if name == "job_equil.rst" else real_read(name)
    with pytest.raises(RuntimeError, match="needs job_equil.rst"):
        runProtocol(_System(), out, _settings(), _hooks([]))


def test_an_interrupted_stage_rewrites_its_trajectory_on_resume(tmp_path):
    """A stride interrupted after some frames left a partial DCD and no
    checkpoint; the resumed run writes that stride's trajectory afresh, so the
    file holds one header and the stride's own frames."""
    from mdcore.dcd import DCD_HEADER_BYTES, readDcd

    class Failing(StubEngine):
        def run(self, count):
            if self.steps >= 14_000:
                raise RuntimeError("interrupted")
            super().run(count)

    out = DirectoryOutputs(str(tmp_path))
    state = {"fail": True}

    def engine_for(which, st):
        cls = Failing if (which == "production" and st.stepCount == 10_000 and state["fail"]) else StubEngine
        return cls(4, 0 if which == "equilibration" else st.stepCount)

    hooks = ProtocolHooks(engineFor=engine_for, initialVelocities=lambda e: np.zeros(12))
    with pytest.raises(RuntimeError, match="interrupted"):
        runProtocol(_System(), out, _settings(), hooks)
    assert out.exists("job_prod_1.dcd") and not out.exists("job_prod_1.rst")
    state["fail"] = False
    rec = runProtocol(_System(), out, _settings(), hooks)
    assert [s["skipped"] for s in rec["stages"]] == [True, False, False]
    data = (tmp_path / "job_prod_1.dcd").read_bytes()
    traj = readDcd(data)
    assert traj.frameCount == 5 and traj.firstStep == 12_000
    assert len(data) == DCD_HEADER_BYTES + 5 * (3 * (8 + 16) + 56)
