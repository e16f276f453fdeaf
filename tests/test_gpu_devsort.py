"""The set-up's device sort and scans (devsort.hip) against numpy: a stable
radix sort of (key, value) pairs by the keys' low end_bit bits, and
exclusive prefix sums, on sizes around the tile boundaries."""
import ctypes as C
import sys

import numpy as np
import pytest

from blood_network_amd import _lib as L

pytestmark = pytest.mark.gpu
lib = L.lib


def case(name):
    print(f"\n[case] {name}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n", [1, 255, 256, 4095, 4096, 4097, 65537, 1_000_003])
def test_exclusive_scans(n):
    case(f"exclusive scans n={n}")
    rng = np.random.default_rng(n)
    a32 = rng.integers(0, 5, n).astype(np.int32)
    a64 = rng.integers(0, 1 << 40, n).astype(np.int64)
    for a, fn in ((a32, lib.bn_dev_scan_i32), (a64, lib.bn_dev_scan_i64)):
        din, dout = L.DeviceArray(a), L.DeviceArray(np.zeros_like(a))
        assert fn(C.c_void_p(din.ptr), C.c_void_p(dout.ptr), n, None) == 0
        ref = np.concatenate([[0], np.cumsum(a, dtype=a.dtype)[:-1]]).astype(a.dtype)
        assert np.array_equal(dout.numpy(), ref)


@pytest.mark.parametrize("n,end_bit", [(1, 8), (300, 13), (4096, 62), (4097, 40), (70001, 17), (2_000_003, 49)])
def test_stable_radix_sort(n, end_bit):
    case(f"stable radix sort n={n} end_bit={end_bit}")
    rng = np.random.default_rng(end_bit)
    # few distinct keys (many ties) and bits above end_bit that must be ignored
    distinct = rng.integers(0, 1 << min(end_bit, 62), max(1, n // 7), dtype=np.uint64)
    keys = distinct[rng.integers(0, len(distinct), n)]
    keys |= rng.integers(0, 2, n, dtype=np.uint64) << np.uint64(63)
    vals = np.arange(n, dtype=np.int32)
    dk, dv = L.DeviceArray(keys), L.DeviceArray(vals)
    ko, vo = L.DeviceArray(np.zeros_like(keys)), L.DeviceArray(np.zeros_like(vals))
    assert lib.bn_dev_sort_pairs(C.c_void_p(dk.ptr), C.c_void_p(ko.ptr), C.c_void_p(dv.ptr), C.c_void_p(vo.ptr), n,
                                 end_bit, None) == 0
    low = keys & np.uint64((1 << end_bit) - 1)
    order = np.argsort(low, kind="stable")
    assert np.array_equal(vo.numpy(), order.astype(np.int32))
    assert np.array_equal(ko.numpy(), keys[order])
    # the input is left as it was
    assert np.array_equal(dk.numpy(), keys) and np.array_equal(dv.numpy(), vals)
