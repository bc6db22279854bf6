"""libssq's device sort and scans (specseqplusplus_amd/csrc/ssq_sort.h), through the test
harness in tests/harness: the stable LSD radix sort against numpy's stable argsort of the same
keys (below 2^end_bit, the sort's contract), and the scans against numpy prefix sums.  Sizes straddle the
4096-key tile and the 256-thread scan block; keys have few distinct values (long runs of
equal digits, where stability shows), one value, descending order and full 64-bit range, with
the bit counts the batched Rips builder sorts by (31 + g and 36 + r bits)."""
import ctypes
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(HERE, "harness", "libssq_sort_harness.so")

SIZES = [0, 1, 2, 255, 256, 257, 4095, 4096, 4097, 3 * 4096 + 17, 4 * 65536 + 3, 1_000_003]


def harness():
    if not os.path.exists(HARNESS):
        raise RuntimeError(f"{HARNESS} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(HARNESS)
    P = ctypes.c_void_p
    L.ssq_harness_sort_u64.argtypes = [P, P, ctypes.c_int64, ctypes.c_int, P, P]
    L.ssq_harness_sort_u64.restype = ctypes.c_int
    L.ssq_harness_scan.argtypes = [P, ctypes.c_int64, ctypes.c_int, P]
    L.ssq_harness_scan.restype = ctypes.c_int
    return L


def test_harness_exports():
    L = harness()
    for name in ("ssq_harness_sort_u64", "ssq_harness_scan"):
        assert hasattr(L, name)


def keys_of(kind, n, rng):
    if kind == "random":
        return rng.integers(0, np.iinfo(np.uint64).max, n, dtype=np.uint64, endpoint=True)
    if kind == "few":
        return rng.integers(0, 4, n).astype(np.uint64) * np.uint64(0x0101010101010101)
    if kind == "one":
        return np.full(n, 0xDEADBEEF12345678, np.uint64)
    if kind == "descending":
        return np.arange(n, 0, -1, dtype=np.uint64) * np.uint64(977)
    raise ValueError(kind)


def sort_on_gpu(L, keys, vals, end_bit):
    n = len(keys)
    keys = np.ascontiguousarray(keys, np.uint64)
    vals = np.ascontiguousarray(vals, np.int32)
    ko = np.zeros(max(n, 1), np.uint64)
    vo = np.zeros(max(n, 1), np.int32)
    rc = L.ssq_harness_sort_u64(keys.ctypes.data, vals.ctypes.data, n, end_bit, ko.ctypes.data, vo.ctypes.data)
    assert rc == 0, f"harness sort failed ({rc})"
    return ko[:n], vo[:n]


@pytest.mark.gpu
@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("kind", ["random", "few", "one", "descending"])
@pytest.mark.parametrize("end_bit", [1, 9, 44, 64])
def test_radix_sort_is_stable_sort(n, kind, end_bit):
    rng = np.random.default_rng(n * 131 + end_bit)
    mask = np.uint64(0xFFFFFFFFFFFFFFFF) if end_bit == 64 else np.uint64((1 << end_bit) - 1)
    keys = keys_of(kind, n, rng) & mask  # the sort's contract: keys below 2^end_bit
    vals = rng.integers(-2**31, 2**31 - 1, n, dtype=np.int64).astype(np.int32)
    print(f"[sort] n={n} kind={kind} end_bit={end_bit}", flush=True)
    ko, vo = sort_on_gpu(harness(), keys, vals, end_bit)
    perm = np.argsort(keys, kind="stable")
    assert np.array_equal(ko, keys[perm])
    assert np.array_equal(vo, vals[perm])


def scan_on_gpu(L, x, mode):
    n = len(x)
    out = np.zeros(n + 1, np.int64)
    rc = L.ssq_harness_scan(x.ctypes.data if n else None, n, mode, out.ctypes.data)
    assert rc == 0, f"harness scan failed ({rc})"
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("n", SIZES + [16 * 4096 * 256 + 5])
def test_scans_equal_prefix_sums(n):
    rng = np.random.default_rng(n)
    L = harness()
    print(f"[scan] n={n}", flush=True)
    a = rng.integers(-1000, 1000, n).astype(np.int32)
    ex = scan_on_gpu(L, a, 0)
    assert ex[0] == 0
    assert np.array_equal(ex[1:], np.cumsum(a, dtype=np.int64))
    u = rng.integers(0, 2**32 - 1, n, dtype=np.int64, endpoint=True).astype(np.uint32)
    ex = scan_on_gpu(L, u, 1)
    assert np.array_equal(ex, np.concatenate([[0], np.cumsum(u, dtype=np.int64)]))
    inc = scan_on_gpu(L, a, 2)[:n]
    assert np.array_equal(inc, np.cumsum(a, dtype=np.int64))
