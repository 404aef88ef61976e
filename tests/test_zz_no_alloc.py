"""mfl_radixsort_compute_noalloc (the sort whose temporaries live in the library's fixed device array)
against mfl_radixsort_compute, both called through the C ABI: bit-identical keys and values, on calls
in a row whose sizes shrink, for one-, two-, three- and four-pass sorts; sizes above the array's limit
are rejected."""
import ctypes

import numpy as np
import pytest

from libcl_amd import _native as N
from libcl_amd import scenes

pytestmark = pytest.mark.gpu

MAX_PAIRS = 1048576  # MFL_SORT_NOALLOC_MAX_PAIRS


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int32).copy()).cuda()


def _run(entry, handle, keys, vals, start, end):
    import torch
    k, v = _dev(keys), _dev(vals)
    ok = getattr(N.load(), entry)(handle, k.data_ptr(), v.data_ptr(), keys.shape[0], start, end, None)
    torch.cuda.synchronize()
    assert ok == 1, N.last_error()
    return k.cpu().numpy().view(np.uint32), v.cpu().numpy().view(np.uint32)


@pytest.fixture(scope="module")
def handles():
    L = N.load()
    a, b = L.mfl_radixsort_create(), L.mfl_radixsort_create()
    assert L.mfl_radixsort_compile(a) == 1 and L.mfl_radixsort_compile(b) == 1, N.last_error()
    yield a, b
    L.mfl_radixsort_destroy(a)
    L.mfl_radixsort_destroy(b)


# (pairs, bit range): calls in this order, each smaller than the one before
CASES = [(MAX_PAIRS, (0, 24)), (300032, (0, 32)), (65536 + 77, (4, 12)), (1000, (0, 16)), (1, (0, 8))]


def test_noalloc_sort_matches_the_allocating_sort(handles):
    plain, noalloc = handles
    for n, (lo, hi) in CASES:
        print(f"noalloc sort n={n} bits [{lo}, {hi})", flush=True)
        keys = scenes.sort_keys(n, seed=n, bits=32 if hi == 32 else 24, skew=(n % 2 == 0))
        vals = np.arange(n, dtype=np.uint32)[::-1].copy()
        ka, va = _run("mfl_radixsort_compute", plain, keys, vals, lo, hi)
        kb, vb = _run("mfl_radixsort_compute_noalloc", noalloc, keys, vals, lo, hi)
        np.testing.assert_array_equal(kb, ka)
        np.testing.assert_array_equal(vb, va)
        digit = (keys.astype(np.uint64) >> np.uint64(lo)) & np.uint64((1 << (hi - lo)) - 1)
        order = np.argsort(digit, kind="stable")
        np.testing.assert_array_equal(kb, keys[order])
        np.testing.assert_array_equal(vb, vals[order])


def test_noalloc_sort_allocates_nothing_and_rejects_large_inputs(handles):
    import torch
    _, noalloc = handles
    L = N.load()
    keys = scenes.sort_keys(MAX_PAIRS, seed=9)
    vals = np.arange(MAX_PAIRS, dtype=np.uint32)
    _run("mfl_radixsort_compute_noalloc", noalloc, keys, vals, 0, 24)
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(3):
        _run("mfl_radixsort_compute_noalloc", noalloc, keys, vals, 0, 24)
    assert abs(torch.cuda.mem_get_info()[0] - free0) < (4 << 20)
    big = torch.zeros(MAX_PAIRS + 256, dtype=torch.int32, device="cuda")
    big_v = torch.zeros_like(big)
    assert L.mfl_radixsort_compute_noalloc(noalloc, big.data_ptr(), big_v.data_ptr(), MAX_PAIRS + 256, 0, 24,
                                           None) == 0
    assert str(MAX_PAIRS) in N.last_error()
    assert L.mfl_radixsort_compute_noalloc(noalloc, big.data_ptr(), big_v.data_ptr(), 1024, 0, 6, None) == 0
    assert "divisible by 4" in N.last_error()
    assert L.mfl_radixsort_compute_noalloc(noalloc, big.data_ptr(), big_v.data_ptr(), 0, 0, 24, None) == 1
