"""The streaming System/State reader (C, one pass, no element tree) against
the element-tree parser it stands in for (CPU).

Every array the engine reads comes out equal, value for value, on the
upstream fixture and on generated systems carrying every accepted force
(chain + water + ions with bonds, angles, torsions, exceptions, restraints,
CMMotionRemover and a barostat, in cubic, octahedral and dodecahedral cells);
the refusals are the same (unknown force, offsets, LJ-PME, switching,
restraint expression and global parameters); and it is fast enough for the
250k-atom configuration (a 250k-atom water box parses in a few seconds).
"""
import os
import time

import numpy as np
import pytest

from conftest import GOLDEN
from mdcore import ingest, synthetic
from mdcore.state import parseState, serialiseState
from mdcore.system import UnsupportedForce, XmlError, parseSystem, serialiseSystem

FIXTURE = open(os.path.join(GOLDEN, "nonbonded-system.xml")).read()


def _same(a, b):
    assert a.particleCount == b.particleCount
    assert a.boxVectors == b.boxVectors
    assert np.array_equal(a.masses, b.masses)
    for name in ("constraintP1", "constraintP2", "constraintDistance"):
        assert np.array_equal(getattr(a, name), getattr(b, name)), name
    for kind in ("HarmonicBondForce", "HarmonicAngleForce", "PeriodicTorsionForce", "NonbondedForce",
                 "PositionRestraint", "CMMotionRemover", "MonteCarloBarostat"):
        fa, fb = a.force(kind), b.force(kind)
        empty = lambda f: f is None or getattr(f, "count", 1) == 0  # noqa: E731
        if empty(fa) or empty(fb):
            assert empty(fa) and empty(fb), kind
            continue
        for field, value in vars(fb).items():
            if field in ("forceGroup", "name", "usesPeriodic", "recipForceGroup", "exceptionsUsePeriodic",
                         "includeDirectSpace", "switchingDistance", "energyExpression", "method", "type"):
                continue
            mine = getattr(fa, field)
            if isinstance(value, np.ndarray):
                assert np.array_equal(mine, value), (kind, field)
            else:
                assert mine == value or tuple(mine) == tuple(value), (kind, field, mine, value)


def test_fixture_reads_the_same():
    _same(ingest.read_system(FIXTURE), parseSystem(FIXTURE))


@pytest.mark.parametrize("shape", ["cube", "octahedron", "dodecahedron"])
def test_generated_systems_with_every_force_read_the_same(shape):
    b = synthetic.build(120, 300, 4, shape, 7, restraint_k=1000.0, barostat=True, grid=(24, 24, 24))
    xml = serialiseSystem(b.system)
    _same(ingest.read_system(xml), parseSystem(xml))
    sx = serialiseState(b.state)
    s1, s2 = ingest.read_state(sx), parseState(sx)
    assert np.array_equal(s1.positions, s2.positions) and np.array_equal(s1.velocities, s2.velocities)
    assert s1.boxVectors == s2.boxVectors and s1.stepCount == s2.stepCount and s1.time == s2.time


def test_state_without_velocities():
    st = parseState(open(os.path.join(GOLDEN, "nonbonded-state.xml")).read())
    st.velocities = None
    s = ingest.read_state(serialiseState(st))
    assert s.velocities is None and np.array_equal(s.positions, st.positions)


@pytest.mark.parametrize("edit, match", [
    (lambda x: x.replace("\t</Forces>", '\t\t<Force forceGroup="0" type="CustomBondForce" version="3"/>\n\t</Forces>'),
     "CustomBondForce"),
    (lambda x: x.replace("<ParticleOffsets/>", '<ParticleOffsets><Offset parameter="l" particle="0" charge="1" '
                                               'sig="0" eps="0"/></ParticleOffsets>'), "offset"),
    (lambda x: x.replace('method="4"', 'method="5"'), "LJ-PME"),
    (lambda x: x.replace('useSwitchingFunction="0"', 'useSwitchingFunction="1"'), "switching"),
])
def test_refusals_match_the_element_tree_parser(edit, match):
    xml = edit(FIXTURE)
    with pytest.raises(UnsupportedForce, match=match):
        ingest.read_system(xml)
    if match != "switching":  # the element-tree parser defers the switching refusal to the engine
        with pytest.raises(UnsupportedForce):
            parseSystem(xml)


def test_restraint_expression_and_globals_are_checked():
    from test_system_contract import RESTRAINT, _with_force
    ok = ingest.read_system(_with_force(RESTRAINT.format(expr="k*periodicdistance(x, y, z, x0, y0, z0)^2;",
                                                         globals="")))
    r = ok.force("PositionRestraint")
    assert list(r.particle) == [8] and r.k[0] == 1000.0 and (r.x0[0], r.y0[0], r.z0[0]) == (1.1, 1.2, 1.3)
    with pytest.raises(UnsupportedForce, match="energy expression"):
        ingest.read_system(_with_force(RESTRAINT.format(expr="0.5*k*periodicdistance(x,y,z,x0,y0,z0)^2",
                                                        globals="")))
    with pytest.raises(UnsupportedForce, match="global parameters"):
        ingest.read_system(_with_force(RESTRAINT.format(expr="k*periodicdistance(x,y,z,x0,y0,z0)^2",
                                                        globals='<Parameter default="1" name="scale"/>')))


def test_malformed_documents():
    with pytest.raises(XmlError):
        ingest.read_system("<State/>")
    with pytest.raises(XmlError):
        ingest.read_system(FIXTURE.replace("<C ", "<D "))
    with pytest.raises(XmlError):
        ingest.read_system(FIXTURE.replace('<Particle mass="14.007"/>', '<Particle mass="14.007"/><Particle mass="1"/>'))


def test_a_quarter_million_atoms_reads_in_seconds():
    b = synthetic.build(0, 83333, 0, "cube", 3)
    xml = serialiseSystem(b.system)
    t = time.perf_counter()
    s = ingest.read_system(xml)
    elapsed = time.perf_counter() - t
    print(f"{s.particleCount} atoms, {len(xml) / 1e6:.1f} MB of XML read in {elapsed:.2f} s")
    assert s.particleCount == b.system.particleCount
    assert elapsed < 10.0
