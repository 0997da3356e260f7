"""The cluster-pair list's capacity on the GPU.

Each super-cluster's entries have a region sized at creation from the
system's density (1.5x the estimate, several times the mean).  A system
compressed far beyond that density must report ListOverflow -- the entries
past a region are never written and every count is clamped to it -- rather
than compute with a truncated list.
"""
import numpy as np
import pytest

from mdcore import native, synthetic

pytestmark = pytest.mark.gpu


def test_compressed_system_reports_list_overflow():
    """The list's capacity is sized from the density the system is created
    at.  Created in a box 1.5 times wider, then compressed to 0.55 of its
    own box (about 20 times the density the list was sized for; the box
    still fits the 1.0 nm cutoff), the search overflows and says so."""
    b = synthetic.config(1)
    box0 = np.asarray(b.system.boxVectors, dtype=np.float64)
    b.system.boxVectors = box0 * 1.5
    s = native.NativeSystem(b.system)
    b.system.boxVectors = box0
    scale = 0.55
    s.set_box(box0 * scale)
    s.set_positions(b.state.positions * scale)
    with pytest.raises(native.MdError) as err:
        s.evaluate()
    assert err.value.name == "ListOverflow", str(err.value)
    s.close()


def test_list_at_the_starting_density_fits():
    b = synthetic.config(1)
    s = native.NativeSystem(b.system)
    s.set_positions(b.state.positions)
    s.evaluate()
    flags = s.list_flags()
    assert not flags["overflowed"]
    assert 0 < flags["tiles"] < flags["tileCapacity"] / 3
    s.close()
