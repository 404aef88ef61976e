"""GPU parity of the key/value radix sort (oclRadixSort::compute, libCL/sort/oclRadixSort.cpp:62-144)
against numpy's stable argsort and the restated reference sort: keys and values bit-exact."""
import numpy as np
import pytest

from libcl_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sorter():
    import libcl_amd as L
    s = L.oclRadixSort()
    assert s.compile() == 1, L.last_error()
    return s


def _expected(keys, vals, start, end):
    w = end - start
    mask = np.uint64((1 << w) - 1)
    digit = (keys.astype(np.uint64) >> np.uint64(start)) & mask
    order = np.argsort(digit, kind="stable")
    return keys[order], vals[order]


def _run(sorter, keys, vals, start, end):
    import torch
    k = torch.from_numpy(keys.view(np.int32).copy()).cuda()
    v = torch.from_numpy(vals.view(np.int32).copy()).cuda()
    ok = sorter.compute(k, v, start, end)
    torch.cuda.synchronize()
    return ok, k.cpu().numpy().view(np.uint32), v.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("n", [1, 7, 64, 255, 256, 257, 1000, 1024, 4099, 65536, 100003, 1 << 20])
# digit widths per pass: (0, 24) 8+8+8, (0, 32) 4 x 8, (4, 12) one 8-bit pass, (0, 20) 7+7+6,
# (8, 28) 7+7+6 above the low byte, (0, 12) 6+6 (widths 7-9 are compiled in, the others run generic)
@pytest.mark.parametrize("bits", [(0, 24), (0, 32), (4, 12), (0, 20), (8, 28), (0, 12)])
def test_sort_matches_stable_argsort(sorter, n, bits):
    print(f"sort n={n} bits={bits}", flush=True)
    keys = scenes.sort_keys(n, seed=100 + n, bits=24) ^ (scenes.splitmix64(7, n) << np.uint64(24)).astype(np.uint32)
    vals = scenes.splitmix64(9, n).astype(np.uint32)
    ok, k, v = _run(sorter, keys, vals, *bits)
    assert ok == 1
    ek, ev = _expected(keys, vals, *bits)
    np.testing.assert_array_equal(k, ek)
    np.testing.assert_array_equal(v, ev)


@pytest.mark.parametrize("case", ["equal", "sorted", "reversed", "two_values", "one_digit_pass"])
def test_sort_special_inputs(sorter, case):
    n = 50000
    vals = np.arange(n, dtype=np.uint32)
    bits = (0, 32)
    if case == "equal":
        keys = np.full(n, 0xABCDEF, np.uint32)
    elif case == "sorted":
        keys = np.arange(n, dtype=np.uint32) * 3
    elif case == "reversed":
        keys = (np.arange(n, dtype=np.uint32) * 3)[::-1].copy()
    elif case == "two_values":
        keys = (scenes.splitmix64(3, n) & np.uint64(1)).astype(np.uint32) * 0xFFFFFFFF
    else:  # one 4-bit pass: a single scatter plus the copy back into the caller's arrays
        keys = scenes.sort_keys(n, seed=5, bits=8)
        bits = (4, 8)
    ok, k, v = _run(sorter, keys, vals, *bits)
    assert ok == 1
    ek, ev = _expected(keys, vals, *bits)
    np.testing.assert_array_equal(k, ek)
    np.testing.assert_array_equal(v, ev)


@pytest.mark.parametrize("skew", [False, True])
def test_sort_16m_cell_keys(sorter, skew):
    """16M uint32 cell-hash keys over 2^24 cells (uniform and 10%-hot-cells skew) with uint32
    particle-index values: bit-exact with a stable sort."""
    n = 1 << 24
    keys = scenes.sort_keys(n, seed=3, skew=skew, bits=24)
    vals = np.arange(n, dtype=np.uint32)
    ok, k, v = _run(sorter, keys, vals, 0, 24)
    assert ok == 1
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(k, keys[order])
    np.testing.assert_array_equal(v, order.astype(np.uint32))


@pytest.mark.parametrize("n", [4096, 1 << 20, (1 << 23) + 5])
def test_sort_unaligned_arrays(sorter, n):
    """Key and value arrays that are 4-byte but not 16-byte aligned (views one element into their
    buffers): whole tiles of aligned arrays are counted with 16-byte loads, these take 4-byte loads."""
    import torch
    print(f"sort unaligned n={n}", flush=True)
    keys = scenes.sort_keys(n, seed=31, bits=24)
    vals = scenes.splitmix64(32, n).astype(np.uint32)
    kb = torch.empty(n + 1, dtype=torch.int32, device="cuda")
    vb = torch.empty(n + 1, dtype=torch.int32, device="cuda")
    k, v = kb[1:], vb[1:]
    assert k.data_ptr() % 16 != 0 and k.is_contiguous()
    k.copy_(torch.from_numpy(keys.view(np.int32)))
    v.copy_(torch.from_numpy(vals.view(np.int32)))
    assert sorter.compute(k, v, 0, 24) == 1
    torch.cuda.synchronize()
    ek, ev = _expected(keys, vals, 0, 24)
    np.testing.assert_array_equal(k.cpu().numpy().view(np.uint32), ek)
    np.testing.assert_array_equal(v.cpu().numpy().view(np.uint32), ev)


def test_sort_reference_smoke(sorter):
    """libCL/TEST/main.cpp:133-181: 1024 rand() keys with identity values, bits 0..32, non-decreasing."""
    keys = (scenes.splitmix64(1, 1024) & np.uint64(0x7FFFFFFF)).astype(np.uint32)
    vals = np.arange(1024, dtype=np.uint32)
    ok, k, v = _run(sorter, keys, vals, 0, 32)
    assert ok == 1
    assert np.all(k[1:] >= k[:-1])
    np.testing.assert_array_equal(keys[v], k)


def test_sort_against_restated_reference(sorter, oracle):
    n = 300000
    keys = scenes.sort_keys(n, seed=21, bits=24)
    vals = np.arange(n, dtype=np.uint32)
    rk, rv = keys.copy(), vals.copy()
    assert oracle.radix_sort(rk, rv, 0, 24)
    ok, k, v = _run(sorter, keys, vals, 0, 24)
    assert ok == 1
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(v, rv)


def test_sort_rejections(sorter):
    import torch
    import libcl_amd as L
    k = torch.zeros(16, dtype=torch.int32, device="cuda")
    v = torch.zeros(16, dtype=torch.int32, device="cuda")
    assert sorter.compute(k, v, 0, 18) == 0  # (end - start) % 4 != 0, oclRadixSort.cpp:76-80
    assert "divisible by 4" in L.last_error()
    assert sorter.compute(k, torch.zeros(8, dtype=torch.int32, device="cuda"), 0, 32) == 0  # size mismatch
    assert sorter.compute(k, v, 8, 8) == 1  # empty bit range is a no-op
    e = torch.zeros(0, dtype=torch.int32, device="cuda")
    assert sorter.compute(e, e.clone(), 0, 32) == 1
