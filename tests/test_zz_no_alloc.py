"""The allocation-free DBSCAN entry point against bifhip_dbscan_device.

bifhip_dbscan_device_noalloc takes every scratch buffer from a fixed
per-GPU device array.  Its counts must equal the allocating entry point's,
bit for bit, on two calls in a row -- the second smaller, so that stale
overflow lists and counters left in the array by the first call would
show -- and on every path: cells the wave kernel holds, cells the
workgroup kernel takes (257..16 384 peaks) and cells beyond it (lane
kernel).  Sizes above the stated limits are refused before any launch.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _peak_sets(rng, sizes, stride):
    nb = len(sizes)
    amps = np.zeros((nb, stride))
    ivls = np.zeros((nb, stride))
    for b, P in enumerate(sizes):
        if P <= 0:
            continue
        if b % 3 == 0:
            a = rng.choice([0.0, 0.02, 0.05], P) + rng.normal(size=P) * 0.002
            iv = rng.integers(1, 3, size=P) * 0.01
        elif b % 3 == 1:
            a = rng.random(P) * 0.5
            iv = rng.random(P) * 0.2
        else:
            a = 0.3 + rng.normal(size=P) * 0.0005
            iv = np.full(P, 0.05)
        amps[b, :P], ivls[b, :P] = a, iv
    return amps, ivls, np.asarray(sizes, dtype=np.int32)


def _both(bh, amps, ivls, cnt, stride, max_peaks):
    nb = len(cnt)
    A = bh.DeviceArray.from_host(amps.ravel(), np.float64)
    I = bh.DeviceArray.from_host(ivls.ravel(), np.float64)
    C = bh.DeviceArray.from_host(cnt, np.int32)
    ref = bh.DeviceArray.full((nb,), -99, np.int32)
    got = bh.DeviceArray.full((nb,), -77, np.int32)
    bh.dbscan_device(A.ptr, I.ptr, stride, nb, C.ptr, max_peaks, 0.01, ref.ptr)
    bh.dbscan_device_noalloc(A.ptr, I.ptr, stride, nb, C.ptr, max_peaks, 0.01, got.ptr)
    bh.synchronize()
    return got.to_host(), ref.to_host()


def test_noalloc_dbscan_equals_allocating_entry_point(orc, bh, dev):
    rng = np.random.default_rng(5)
    # call 1: every path, rows of 16 400 slots (the lane kernel takes 16 390)
    sizes1 = [0, 1, 2, 40, 85, 256, 257, 900, 3073, 6000, 16384, 16390, -1, 120, 16390, 300]
    stride1 = 16400
    amps, ivls, cnt = _peak_sets(rng, sizes1, stride1)
    say("noalloc call 1", len(sizes1), "blocks, stride", stride1)
    got, ref = _both(bh, amps, ivls, cnt, stride1, 1 << 30)
    np.testing.assert_array_equal(got, ref)
    # call 2: fewer, shorter rows; overflow cells again, but fewer of them
    sizes2 = [3, 300, 2, 257, 64, 65]
    stride2 = 320
    amps, ivls, cnt = _peak_sets(rng, sizes2, stride2)
    say("noalloc call 2", len(sizes2), "blocks, stride", stride2)
    got2, ref2 = _both(bh, amps, ivls, cnt, stride2, 1 << 30)
    np.testing.assert_array_equal(got2, ref2)
    for b, P in enumerate(sizes2):
        assert got2[b] == orc.dbscan(amps[b, :P], ivls[b, :P], 0.01), (b, P)
    # call 3: a wave-only batch with maxAmountOfPeaks below some counts
    rng2 = np.random.default_rng(9)
    sizes3 = list(rng2.integers(1, 200, size=500))
    amps, ivls, cnt = _peak_sets(rng2, sizes3, 200)
    say("noalloc call 3", len(sizes3), "blocks")
    got3, ref3 = _both(bh, amps, ivls, cnt, 200, 120)
    np.testing.assert_array_equal(got3, ref3)


def test_noalloc_dbscan_refuses_sizes_beyond_its_limits(bh, dev):
    A = bh.DeviceArray((16,), np.float64).zero_()
    C = bh.DeviceArray((4,), np.int32).zero_()
    out = bh.DeviceArray((4,), np.int32).zero_()
    with pytest.raises(bh.BifhipError) as e:
        bh.dbscan_device_noalloc(A.ptr, A.ptr, 4, bh.NOALLOC_MAX_BLOCKS + 1, C.ptr, 10, 0.01,
                                 out.ptr)
    assert e.value.code == bh.E_LIMIT
    with pytest.raises(bh.BifhipError) as e:
        bh.dbscan_device_noalloc(A.ptr, A.ptr, bh.NOALLOC_MAX_ROW + 1, 1, C.ptr, 10, 0.01,
                                 out.ptr)
    assert e.value.code == bh.E_LIMIT
