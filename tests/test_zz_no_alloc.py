"""md_probe_fixed_point_noalloc: the fixed-point probe with its staging in a
fixed device array instead of per-call allocations (for callers that cannot
allocate).  Through the C ABI, on two calls in a row -- the second with a
smaller input, so stale staging from the first would show -- its outputs
equal md_probe_fixed_point's bit for bit; sizes past its maximum are refused
by name."""
import ctypes as C

import numpy as np
import pytest

from mdcore import native

pytestmark = pytest.mark.gpu

MAX_VALUES = 65536


def _call(name, values):
    L = native.library()
    fn = getattr(L, name)
    fn.restype = C.c_int
    n = values.shape[0]
    fixed = np.zeros(n, dtype=np.int64)
    back = np.zeros(n, dtype=np.float32)
    rc = fn(values.ctypes.data_as(C.POINTER(C.c_float)), C.c_int32(n), fixed.ctypes.data_as(C.POINTER(C.c_int64)),
            back.ctypes.data_as(C.POINTER(C.c_float)))
    return rc, fixed, back


def test_no_alloc_probe_matches_the_allocating_probe_twice():
    rng = np.random.default_rng(5)
    first = (rng.uniform(-1, 1, 50000) * 10.0 ** rng.uniform(-9, 6, 50000)).astype(np.float32)
    second = (rng.uniform(-1, 1, 777) * 10.0 ** rng.uniform(-9, 6, 777)).astype(np.float32)
    for values in (first, second):
        rc0, f0, b0 = _call("md_probe_fixed_point", values)
        rc1, f1, b1 = _call("md_probe_fixed_point_noalloc", values)
        assert rc0 == 0 and rc1 == 0, native.library().md_last_error()
        assert np.array_equal(f0, f1)
        assert np.array_equal(b0, b1)


def test_no_alloc_probe_refuses_more_than_its_staging_holds():
    rc, _, _ = _call("md_probe_fixed_point_noalloc", np.zeros(MAX_VALUES + 1, dtype=np.float32))
    assert native.ERRORS[rc] == "InvalidArgument"
    assert b"at most" in native.library().md_last_error()
    rc, _, _ = _call("md_probe_fixed_point_noalloc", np.zeros(MAX_VALUES, dtype=np.float32))
    assert rc == 0
