"""gc_SortNoAlloc, the gc_Sort variant that allocates no device memory (its
scratch is a fixed per-GPU array): bit-identical to gc_Sort and to the
oracle's (key, value) sort on two calls in a row, the second smaller than
the first (what the first left in the array must not leak into it); sizes
beyond its limits are refused before any kernel runs."""
import numpy as np
import pytest

from gpucell_amd import abi
from oracle import pyoracle as orc
from tests import gpu_util as gu

pytestmark = pytest.mark.gpu

MAX_N = MAX_K = 1 << 22


def _sort(name, keys, vals):
    k, v = gu.dev(keys, np.int32), gu.dev(vals, np.int32)
    gu.call(name, int(keys.shape[0]), k, v)
    return gu.host(k), gu.host(v)


def test_noalloc_sort_matches_sort_on_two_calls(gpu):
    rng = np.random.default_rng(11)
    inputs = []
    for n, K in ((600_000, 150_000), (25_000, 700)):   # the second call is smaller
        keys = rng.integers(0, K, n).astype(np.int32)
        keys[rng.integers(0, n, 50)] = K - 1              # the largest key is present
        vals = rng.permutation(n).astype(np.int32)
        inputs.append((keys, vals))
    refs = [_sort("gc_Sort", k, v) for k, v in inputs]
    outs = [_sort("gc_SortNoAlloc", k, v) for k, v in inputs]   # two calls in a row
    for (keys, vals), (kr, vr), (kn, vn) in zip(inputs, refs, outs):
        assert np.array_equal(kn, kr) and np.array_equal(vn, vr)
        ko, vo = orc.sort(keys, vals)
        assert np.array_equal(kn, ko) and np.array_equal(vn, vo)


def test_noalloc_sort_refuses_sizes_beyond_its_limits(gpu):
    lib = abi.load()
    small = gu.dev(np.zeros(16), np.int32)
    # too many elements: refused before any kernel reads the arrays
    st = lib.gc_SortNoAlloc(MAX_N + 1, abi.ptr(small), abi.ptr(small))
    assert st == abi.GC_ERR_INVALID and b"GC_NOALLOC_MAX_N" in lib.gc_GetLastError()
    # a key at the limit: refused after the key-range pass
    keys = gu.dev(np.array([3, MAX_K, 7, 1], np.int32), np.int32)
    vals = gu.dev(np.arange(4), np.int32)
    st = lib.gc_SortNoAlloc(4, abi.ptr(keys), abi.ptr(vals))
    assert st == abi.GC_ERR_INVALID and b"GC_NOALLOC_MAX_K" in lib.gc_GetLastError()
    assert np.array_equal(gu.host(keys), [3, MAX_K, 7, 1])   # untouched
    # an over-full key is refused as by gc_Sort
    n = 9000
    keys = gu.dev(np.zeros(n), np.int32)
    vals = gu.dev(np.arange(n), np.int32)
    assert lib.gc_SortNoAlloc(n, abi.ptr(keys), abi.ptr(vals)) == abi.GC_ERR_GRID
    # and a legal call afterwards still sorts
    k, v = _sort("gc_SortNoAlloc", np.array([5, 2, 5, 0], np.int32), np.array([3, 1, 0, 2], np.int32))
    assert np.array_equal(k, [0, 2, 5, 5]) and np.array_equal(v, [2, 1, 0, 3])
