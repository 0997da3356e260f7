"""The System/State input contract (CPU): what is accepted, what is refused,
and exact serialise/parse round trips.

Mirrors ``src/engine/system.ts`` (``UnsupportedForce`` ``:178-190``, the
offset and LJ-PME refusals ``:371-380``, the restraint expression check
``:402-412``) and ``test/state-serialise.test.ts`` (exact round trips with
awkward values; velocities absent rather than invented; forces only when
asked).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from mdcore.state import OpenMMState, parseState, serialiseState
from mdcore.system import UnsupportedForce, XmlError, parseSystem, serialiseSystem

FIXTURE = open(os.path.join(GOLDEN, "nonbonded-system.xml")).read()

RESTRAINT = """
		<Force energy="{expr}" forceGroup="2" name="PositionRestraint" type="CustomExternalForce" version="3">
			<PerParticleParameters>
				<Parameter name="k"/>
				<Parameter name="x0"/>
				<Parameter name="y0"/>
				<Parameter name="z0"/>
			</PerParticleParameters>
			<GlobalParameters>{globals}</GlobalParameters>
			<EnergyParameterDerivatives/>
			<Particles>
				<Particle index="8" param1="1000" param2="1.1" param3="1.2" param4="1.3"/>
			</Particles>
		</Force>
"""


def _with_force(xml_force: str) -> str:
    return FIXTURE.replace("\t</Forces>", xml_force + "\t</Forces>")


def test_fixture_parses_and_round_trips_exactly():
    s = parseSystem(FIXTURE)
    assert s.particleCount == 10
    t = parseSystem(serialiseSystem(s))
    assert np.array_equal(t.masses, s.masses)
    a, b = s.force("NonbondedForce"), t.force("NonbondedForce")
    for name in ("charge", "sigma", "epsilon", "exceptionP1", "exceptionP2", "exceptionChargeProd",
                 "exceptionSigma", "exceptionEpsilon"):
        assert np.array_equal(getattr(a, name), getattr(b, name)), name
    assert (a.cutoff, a.ewaldTolerance, a.method, a.useDispersionCorrection) == \
        (b.cutoff, b.ewaldTolerance, b.method, b.useDispersionCorrection)
    assert t.boxVectors == s.boxVectors


def test_an_unknown_force_is_refused_by_name():
    xml = _with_force('\t\t<Force forceGroup="0" name="X" type="CustomBondForce" version="3"/>\n')
    with pytest.raises(UnsupportedForce, match="CustomBondForce"):
        parseSystem(xml)


def test_parameter_offsets_are_refused():
    xml = FIXTURE.replace("<ParticleOffsets/>",
                          '<ParticleOffsets><Offset charge="1" eps="0" parameter="lambda" particle="0" sig="0"/>'
                          "</ParticleOffsets>")
    with pytest.raises(UnsupportedForce, match="offset"):
        parseSystem(xml)


def test_lj_pme_is_refused():
    xml = FIXTURE.replace('method="4"', 'method="5"')
    with pytest.raises(UnsupportedForce, match="LJ-PME"):
        parseSystem(xml)


def test_the_restraint_expression_is_checked():
    ok = parseSystem(_with_force(RESTRAINT.format(expr="k*periodicdistance(x, y, z, x0, y0, z0)^2;", globals="")))
    r = ok.force("PositionRestraint")
    assert r is not None and list(r.particle) == [8] and r.k[0] == 1000.0 and r.x0[0] == 1.1
    with pytest.raises(UnsupportedForce, match="energy expression"):
        parseSystem(_with_force(RESTRAINT.format(expr="0.5*k*periodicdistance(x,y,z,x0,y0,z0)^2", globals="")))
    with pytest.raises(UnsupportedForce, match="global parameters"):
        parseSystem(_with_force(RESTRAINT.format(expr="k*periodicdistance(x,y,z,x0,y0,z0)^2",
                                                 globals='<Parameter default="1" name="scale"/>')))


def test_particle_count_mismatch_is_an_xml_error():
    xml = FIXTURE.replace('<Particle mass="14.007"/>', '<Particle mass="14.007"/><Particle mass="1.0"/>', 1)
    with pytest.raises(XmlError, match="particles"):
        parseSystem(xml)


BOX = ((3.080639849905961, 0.0, 0.0), (0.0, 3.080639849905961, 0.0), (0.0, 0.0, 3.080639849905961))
MADE = OpenMMState(
    boxVectors=BOX,
    positions=np.array([-1.2345678901234567, 0, 3.141592653589793, 1e-12, 2.5, -0.000123456789, 6.02e-7,
                        1.9999999999999998, 0]),
    velocities=np.array([-0.5000000000000004, 0.5000000000000004, 0, 0.24999999999997247, -0.24999999999997247,
                         0.12500000000004174, 1.5, -1.5, 0.75]),
    stepCount=12500, time=25.000000000000004)


def test_state_round_trips_every_value_exactly():
    p = parseState(serialiseState(MADE))
    assert p.stepCount == MADE.stepCount and p.time == MADE.time
    assert p.particleCount == 3
    assert np.array_equal(p.positions, MADE.positions)
    assert np.array_equal(p.velocities, MADE.velocities)
    assert p.boxVectors == MADE.boxVectors


def test_state_without_velocities_reads_back_without_them():
    s = OpenMMState(boxVectors=BOX, positions=MADE.positions)
    xml = serialiseState(s)
    assert "<Velocities>" not in xml
    assert parseState(xml).velocities is None


def test_forces_are_written_only_when_asked():
    s = OpenMMState(boxVectors=BOX, positions=MADE.positions, velocities=MADE.velocities,
                    forces=np.arange(1.0, 10.0))
    assert "<Forces>" not in serialiseState(s)
    xml = serialiseState(s, includeForces=True)
    assert list(parseState(xml).forces) == list(range(1, 10))


def test_state_fixture_round_trips():
    text = open(os.path.join(GOLDEN, "nonbonded-state.xml")).read()
    s = parseState(text)
    t = parseState(serialiseState(s, includeForces=True))
    assert np.array_equal(t.positions, s.positions)
    assert t.boxVectors == s.boxVectors
