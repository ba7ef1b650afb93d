"""fsx_indexx_noalloc: the exact edge sort with no device allocation (working arrays in a fixed area of
the library's device memory, one per GPU), against fsx_indexx (which allocates a context for them):
the same 1-based permutation on two calls in a row, the second smaller, and n above the area's
limit refused."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _keys(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, max(2, n // 5), n).astype(np.float64) * 0.25
    a[rng.random(n) < 0.1] = -1e10  # the virtual edges' weight
    return a


def test_indexx_noalloc_matches_indexx():
    from fastscapelib_fortran import _lib
    lib = _lib.lib
    for n, seed in [(300000, 1), (4999, 2)]:
        keys = _keys(n, seed)
        want = _lib.indexx(keys)
        kd = _lib.DeviceBuffer(8 * n).upload(keys)
        pd = _lib.DeviceBuffer(4 * n).upload(np.full(n, -7, np.int32))
        print(f"indexx_noalloc: n={n}", flush=True)
        _lib.check(lib.fsx_indexx_noalloc(n, kd.ptr, pd.ptr, None, 1), "fsx_indexx_noalloc")
        got = pd.download(np.empty(n, np.int32))
        kd.free()
        pd.free()
        assert np.array_equal(got, want), n
        assert np.all(np.diff(keys[got - 1]) >= 0)
    big = _lib.FSX_INDEXX_NOALLOC_MAX + 1
    assert lib.fsx_indexx_noalloc(big, C.c_void_p(1), C.c_void_p(1), None, 1) == _lib.FSX_EINVAL
    assert lib.fsx_indexx_noalloc(0, None, None, None, 1) == 0
