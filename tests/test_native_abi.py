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
