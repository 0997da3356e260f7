"""The C ABI library on the CPU lane: it loads, exports every entry point the
header declares, and fails loudly (no silent fallback) where there is no GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from mdcore import native, synthetic

HEADER = os.path.join(ROOT, "include", "mdcore.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(md_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(native.LIB_PATH)
    names = declared_symbols()
    assert "md_system_create" in names and "md_run" in names
    for name in names:
        assert hasattr(lib, name), name


def test_version_and_empty_error():
    L = native.library()
    assert b"gfx950" in L.md_version()


def test_create_without_a_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present; the GPU lane covers creation")
    b = synthetic.config(0)
    with pytest.raises(native.NoDevice):
        native.NativeSystem(b.system)


def test_descriptor_packs_every_force():
    b = synthetic.build(40, 30, 2, "cube", 3, restraint_k=1000.0, barostat=True)
    d, keep = native.describe(b.system)
    s = b.system
    assert d.n_particles == s.particleCount
    assert d.n_constraints == s.constraintCount
    assert d.n_bonds == s.force("HarmonicBondForce").count
    assert d.n_exceptions == s.force("NonbondedForce").exceptionCount
    assert d.n_restraints > 0 and d.has_barostat == 1 and d.cmm_frequency == 1
    assert abs(d.cutoff - 1.0) < 1e-15


@pytest.mark.parametrize("dims", [(2, 2, 2), (8, 8, 8), (22, 22, 22), (8, 10, 14), (4, 2, 30), (20, 18, 16),
                                  (2, 2, 64), (30, 4, 2)])
def test_cell_walk_visits_every_cell_through_face_neighbours(dims):
    """The cell sort order (host side of the library, no device): each cell
    once, each step to a face neighbour -- runs of cells never jump across
    the grid or wrap the box."""
    cells = native.probe_cell_order(dims)
    W, H, D = dims
    assert np.array_equal(np.sort(cells), np.arange(W * H * D))
    x, y, z = cells // (H * D), (cells // D) % H, cells % D
    step = np.abs(np.diff(x)) + np.abs(np.diff(y)) + np.abs(np.diff(z))
    assert np.all(step == 1)


def test_cell_walk_rejects_bad_grids():
    with pytest.raises(native.MdError):
        native.probe_cell_order((0, 4, 4))


# ---- named errors the C ABI raises before it touches a device ------------------
# md_system_create validates and packs the whole System on the host first, so
# each refusal below is the same with or without a GPU (upstream's named
# errors: ClusterTooLarge constraint-clusters.ts:71-83, CutoffTooLargeForBox
# triclinic.ts:179, NotSmooth fft.ts:24, UnsupportedForce system.ts:178).

def _tiny(n_atoms=4, box=3.0):
    from mdcore.system import CMMotionRemover, NonbondedForce, OpenMMSystem
    nb = NonbondedForce(charge=np.zeros(n_atoms), sigma=np.full(n_atoms, 0.3), epsilon=np.full(n_atoms, 0.5),
                        exceptionP1=np.zeros(0, np.int32), exceptionP2=np.zeros(0, np.int32),
                        exceptionChargeProd=np.zeros(0), exceptionSigma=np.zeros(0), exceptionEpsilon=np.zeros(0),
                        cutoff=1.0, ewaldTolerance=5e-4, grid=(0, 0, 0), useDispersionCorrection=False)
    bv = ((box, 0.0, 0.0), (0.0, box, 0.0), (0.0, 0.0, box))
    return OpenMMSystem(boxVectors=bv, masses=np.full(n_atoms, 12.0), constraintP1=np.zeros(0, np.int32),
                        constraintP2=np.zeros(0, np.int32), constraintDistance=np.zeros(0),
                        forces=[nb, CMMotionRemover(1)])


def test_a_constraint_cluster_beyond_eight_atoms_is_refused():
    import dataclasses
    s = _tiny(10)
    # a chain of nine constraints joins ten atoms into one cluster
    s = dataclasses.replace(s, constraintP1=np.arange(9, dtype=np.int32),
                            constraintP2=np.arange(1, 10, dtype=np.int32), constraintDistance=np.full(9, 0.15))
    with pytest.raises(native.ClusterTooLarge):
        native.NativeSystem(s)


def test_a_cutoff_past_half_the_box_is_refused():
    with pytest.raises(native.CutoffTooLargeForBox):
        native.NativeSystem(_tiny(4, box=1.9))


def test_a_mesh_with_a_large_prime_factor_is_refused():
    with pytest.raises(native.MdError) as err:
        native.NativeSystem(_tiny(4), mesh=(34, 32, 32))  # 34 = 2 * 17
    assert err.value.name == "NotSmooth"


def test_a_switching_function_is_an_unsupported_force():
    import dataclasses
    s = _tiny(4)
    nb = dataclasses.replace(s.force("NonbondedForce"), useSwitchingFunction=True, switchingDistance=0.9)
    s = dataclasses.replace(s, forces=[nb])
    with pytest.raises(native.MdError) as err:
        native.NativeSystem(s)
    assert err.value.name == "UnsupportedForce"


def test_an_out_of_range_index_is_an_invalid_argument():
    import dataclasses
    s = dataclasses.replace(_tiny(4), constraintP1=np.array([0], np.int32), constraintP2=np.array([7], np.int32),
                            constraintDistance=np.array([0.1]))
    with pytest.raises(native.MdError) as err:
        native.NativeSystem(s)
    assert err.value.name == "InvalidArgument"
