"""Monte Carlo barostat host logic (src/reference/barostat.ts,
src/engine/molecules.ts) on the CPU, with the oracle's potential standing in
for the device's.

Pins: the molecule count of a generated system is known by construction
(one chain, each water, each ion); centroid scaling is exact and leaves every
intramolecular distance alone; the counter fires on the frequency-th call and
carries overshoot (barostat.ts:118-139); the Metropolis work of a
constant-energy move has the closed form P dV - N kT ln(V'/V); the proposal
width adapts by 1.1 after ten attempts and is capped at 0.3 V.
"""
import math

import numpy as np
import pytest

import oracle as O
from mdcore import barostat as baro
from mdcore import synthetic
from mdcore.units import BOLTZMANN_KJ_PER_MOL_K


@pytest.fixture(scope="module")
def small():
    return synthetic.build(40, 30, 2, "cube", 3, barostat=True)


def test_molecules_are_chain_waters_and_ions(small):
    m = baro.find_molecules(small.system)
    assert m.count == 1 + 30 + 2
    assert m.largest == 40
    # numbered by lowest atom: the chain (atoms 0..39) is molecule 0
    assert np.all(m.moleculeOf[:40] == 0)
    assert m.starts[-1] == small.system.particleCount
    # atoms grouped by molecule, ascending inside each group
    for k in range(m.count):
        grp = m.atoms[m.starts[k]:m.starts[k + 1]]
        assert np.all(np.diff(grp) > 0)
        assert np.all(m.moleculeOf[grp] == k)


def test_scaling_is_rigid_per_molecule(small):
    s = small.system
    m = baro.find_molecules(s)
    x = small.state.positions.reshape(-1, 3)
    y = baro.scale_molecules(small.state.positions, m, 1.013).reshape(-1, 3)
    r0 = np.linalg.norm(x[s.constraintP1] - x[s.constraintP2], axis=1)
    r1 = np.linalg.norm(y[s.constraintP1] - y[s.constraintP2], axis=1)
    assert np.max(np.abs(r1 - r0)) < 1e-12
    for k in (0, 5, m.count - 1):
        grp = m.atoms[m.starts[k]:m.starts[k + 1]]
        np.testing.assert_allclose(y[grp].mean(axis=0), 1.013 * x[grp].mean(axis=0), rtol=0, atol=1e-12)
        d = y[grp] - x[grp]
        assert np.max(np.abs(d - d[0])) < 1e-12  # one translation per molecule


def test_counter_fires_on_the_frequency_th_call():
    st = baro.BarostatState(step=0, volumeScale=1.0)
    fired = [baro.acts_after_steps(st, 25, 1) for _ in range(50)]
    assert [i + 1 for i, f in enumerate(fired) if f] == [25, 50]
    st = baro.BarostatState(step=0, volumeScale=1.0)
    assert not baro.acts_after_steps(st, 25, 10)
    assert baro.acts_after_steps(st, 25, 20)
    assert st.step == 5  # overshoot carries
    assert not baro.acts_after_steps(st, 0, 100)


class _Baro:
    defaultPressure = 1.0
    defaultTemperature = 300.0
    frequency = 25


def test_work_of_a_constant_energy_move_has_its_closed_form(small):
    m = baro.find_molecules(small.system)
    box = np.asarray(small.system.boxVectors)
    st = baro.BarostatState.initial(box)
    draws = iter([0.75, 0.999999])
    res = baro.attempt(_Baro, st, box, small.state.positions, lambda x, b: 123.0, lambda: next(draws), m)
    v = baro.box_volume(box)
    dv = 0.01 * v * 2 * (0.75 - 0.5)
    kT = BOLTZMANN_KJ_PER_MOL_K * 300.0
    expected = 1.0 * 6.02214076e23 * 1e-25 * dv - m.count * kT * math.log((v + dv) / v)
    assert res.deltaVolume == pytest.approx(dv, rel=1e-15)
    assert res.work == pytest.approx(expected, rel=1e-12, abs=1e-12)
    assert res.lengthScale == pytest.approx(((v + dv) / v) ** (1 / 3), rel=1e-15)
    # an expanding move at 1 bar is downhill here (the ln term wins): no second draw
    assert (res.work <= 0) == res.accepted
    np.testing.assert_allclose(res.box, box * res.lengthScale if res.accepted else box)


def test_uphill_move_is_rejected_by_a_large_deviate_and_width_adapts(small):
    m = baro.find_molecules(small.system)
    box = np.asarray(small.system.boxVectors)
    st = baro.BarostatState.initial(box)
    w0 = st.volumeScale
    # energy rises steeply with volume: every move up is uphill
    energy = lambda x, b: 1e6 * baro.box_volume(b)
    for _ in range(10):
        draws = iter([0.9, 0.999])
        res = baro.attempt(_Baro, st, box, small.state.positions, energy, lambda: next(draws), m)
        assert not res.accepted
        np.testing.assert_array_equal(res.positions, small.state.positions)
    assert st.volumeScale == pytest.approx(w0 / 1.1)
    assert st.attempted == 0 and st.accepted == 0
    # all accepted: grows by 1.1, capped at 0.3 V
    st = baro.BarostatState(step=0, volumeScale=0.29 * baro.box_volume(box))
    for _ in range(10):
        draws = iter([0.1, 0.0])
        baro.attempt(_Baro, st, box, small.state.positions, lambda x, b: 0.0, lambda: next(draws), m)
    assert st.volumeScale == pytest.approx(0.3 * baro.box_volume(box))


def test_attempt_on_the_oracle_potential(small):
    """A real move with the f64 oracle as the potential: the accepted
    configuration keeps every constraint and the work adds up."""
    s = small.system
    m = baro.find_molecules(s)
    o = O.OracleSystem(s)
    box = np.asarray(s.boxVectors)
    grid = o.default_grid(box)

    def potential(x, b):
        total, _, _ = o.potential(np.asarray(x), b, grid=grid)
        return total

    st = baro.BarostatState.initial(box)
    draws = iter([0.3, 0.5])
    res = baro.attempt(_Baro, st, box, small.state.positions, potential, lambda: next(draws), m)
    assert np.isfinite(res.work)
    kT = BOLTZMANN_KJ_PER_MOL_K * 300.0
    v = baro.box_volume(box)
    assert res.work == pytest.approx(res.finalEnergy - res.initialEnergy
                                     + 6.02214076e23 * 1e-25 * res.deltaVolume
                                     - m.count * kT * math.log((v + res.deltaVolume) / v), rel=1e-12)
    assert res.initialEnergy == pytest.approx(potential(small.state.positions, box), rel=1e-14)


def test_molecule_table_is_what_the_device_barostat_takes(small):
    """md_barostat_attach takes molecules as a CSR table: starts from 0 to
    N, every particle exactly once, ascending within each molecule (the
    order the device sums a centroid in, as numpy's bincount does)."""
    mol = baro.find_molecules(small.system)
    n = small.system.particleCount
    assert mol.starts[0] == 0 and mol.starts[-1] == n
    assert np.all(np.diff(mol.starts) > 0)
    assert np.array_equal(np.sort(mol.atoms), np.arange(n))
    for m in range(mol.count):
        block = mol.atoms[mol.starts[m]:mol.starts[m + 1]]
        assert np.all(np.diff(block) > 0)
        assert np.all(mol.moleculeOf[block] == m)
