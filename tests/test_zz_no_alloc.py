"""The allocation-free sort (mpm_sort_pairs_host_noalloc) against the allocating
one (mpm_sort_pairs_host), through the C-ABI (GPU).

Two calls in a row, the second with fewer keys than the first (the device array
keeps the first call's contents), then its limits: the maximum key count is
accepted, one more key is refused with MPM_E_INVALID and leaves the library
usable, and an empty input is a no-op.  Sorted keys and permutations are integer
results: bit-exact against the allocating entry point and against NumPy's stable
argsort."""
import ctypes

import numpy as np
import pytest

from mpm_snow_sim2d_amd import _lib

pytestmark = pytest.mark.gpu

MAXK = 1 << 20


def _sort(fn, keys, key_bits):
    keys = np.ascontiguousarray(keys, np.uint32)
    ok = np.empty(keys.size, np.uint32)
    op = np.empty(keys.size, np.uint32)
    vp = ctypes.c_void_p
    r = fn(keys.ctypes.data_as(vp), keys.size, key_bits, ok.ctypes.data_as(vp), op.ctypes.data_as(vp))
    return r, ok, op


def test_noalloc_sort_matches_allocating_sort():
    lib = _lib.load()
    rng = np.random.default_rng(0)
    cases = [(rng.integers(0, 1 << 25, 300_000, dtype=np.uint64).astype(np.uint32), 25),
             (rng.integers(0, 1 << 9, 1000, dtype=np.uint64).astype(np.uint32), 9),     # smaller, long runs
             (rng.integers(0, 1 << 32, MAXK, dtype=np.uint64).astype(np.uint32), 32)]   # the maximum
    for keys, bits in cases:
        print(f"[noalloc] n={keys.size} key_bits={bits}", flush=True)
        r0, k0, p0 = _sort(lib.mpm_sort_pairs_host, keys, bits)
        r1, k1, p1 = _sort(lib.mpm_sort_pairs_host_noalloc, keys, bits)
        assert r0 == 0 and r1 == 0, lib.mpm_last_error()
        ref = np.argsort(keys, kind="stable").astype(np.uint32)
        assert np.array_equal(p1, p0) and np.array_equal(k1, k0)
        assert np.array_equal(p1, ref) and np.array_equal(k1, keys[ref])


def test_noalloc_sort_limits():
    lib = _lib.load()
    too_many = np.zeros(MAXK + 1, np.uint32)
    r, _, _ = _sort(lib.mpm_sort_pairs_host_noalloc, too_many, 8)
    assert r == _lib.MPM_E_INVALID
    assert b"exceeds" in lib.mpm_last_error()
    r, _, _ = _sort(lib.mpm_sort_pairs_host_noalloc, np.zeros(0, np.uint32), 8)
    assert r == 0
    r, _, _ = _sort(lib.mpm_sort_pairs_host_noalloc, np.array([300], np.uint32), 8)   # key beyond key_bits
    assert r == _lib.MPM_E_INVALID
    keys = np.array([5, 3, 5, 1, 3, 0], np.uint32)
    r, k, p = _sort(lib.mpm_sort_pairs_host_noalloc, keys, 3)
    assert r == 0 and p.tolist() == [5, 3, 1, 4, 0, 2] and k.tolist() == [0, 1, 3, 3, 5, 5]
