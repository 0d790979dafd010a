"""The allocation-free binning entry point (mcm_downsample_params_noalloc)
against mcm_downsample_params through the C ABI: bit for bit, on two calls in
a row (the second smaller, so stale contents of the shared scratch would
show), and sizes beyond the scratch are errors."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from mcmatlab_amd import _lib

    L = _lib.lib()
    assert L.mcm_device_count() >= 1
    return L


def call(L, name, pv, n2, *extra):
    pv = np.ascontiguousarray(np.asarray(pv, np.float32).reshape(3, -1).T.ravel())  # interleaved sets, column-major 3 x n1
    n1 = pv.size // 3
    out = np.zeros(3 * n2, np.float32)
    bins = np.zeros(n1, np.uint8)
    rc = getattr(L, name)(pv.ctypes.data, n1, n2, out.ctypes.data, bins.ctypes.data, 0, *extra)
    return rc, out, bins


def test_noalloc_equals_allocating_entry_point(lib):
    rng = np.random.default_rng(17)
    for n1, n2, kind in ((300000, 60, "distinct"), (70001, 255, "ties"), (5000, 300, "distinct")):
        pv = rng.random((3, n1)) if kind == "distinct" else rng.integers(0, 5, (3, n1))
        pv = pv.astype(np.float32)
        print(f"[no-alloc] n1={n1} n2={n2} {kind}", flush=True)
        rc0, o0, b0 = call(lib, "mcm_downsample_params", pv, n2)
        rc1, o1, b1 = call(lib, "mcm_downsample_params_noalloc", pv, n2, None)
        assert rc0 == 0 and rc1 == 0
        assert np.array_equal(o0.view(np.uint32), o1.view(np.uint32))
        assert np.array_equal(b0, b1)


def test_noalloc_rejects_sizes_beyond_its_scratch(lib):
    from mcmatlab_amd import abi

    pv = np.zeros((3, 4), np.float32)
    rc, _, _ = call(lib, "mcm_downsample_params_noalloc", pv, 5000, None)
    assert rc == abi.MCM_ERR_INVALID_ARGUMENT
    assert b"bins" in lib.mcm_last_error()
    big = np.zeros((3, (1 << 20) + 1), np.float32)
    rc, _, _ = call(lib, "mcm_downsample_params_noalloc", big, 4, None)
    assert rc == abi.MCM_ERR_INVALID_ARGUMENT
