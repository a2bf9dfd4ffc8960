"""GPU test of the fused LLE + peak-count map.

The fused kernel must reproduce, bit for bit, both the LLE2D map and the
bifurcation2DDBSCAN map of the same grid (each already pinned against the
oracle in test_gpu_parity.py), and the oracle directly on sampled rows.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

IC = [0.1, 0.1, 0.1]
VALS = [0.2, 0.2, 5.7]
R2 = [0.05, 0.45, 2.0, 10.0]


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)],
                                                                       b[~np.isnan(b)])


@pytest.mark.parametrize("pert", [0, 1])
def test_fused_equals_separate_maps(bh, orc, pert):
    n = 48
    # peak window [20, 60) t.u. lies inside the LLE trajectory 10 + 50 t.u.
    peak = dict(tmax=40.0, trans=20.0)
    lle = dict(tmax=50.0, NT=0.5, trans=10.0, eps=1e-8)
    bh.set_perturbation(pert)
    say("fused vs separate", n, pert)
    d_peak = bh.DeviceArray((n * n,), np.int32).zero_()
    d_lle = bh.DeviceArray((n * n,), np.float64).zero_()
    bh.lle_peak_map_device(d_peak.ptr, d_lle.ptr, n, 0.01, IC, R2, [0, 2], VALS, peak["tmax"],
                           peak["trans"], 0, 100000, 0.01, lle["tmax"], lle["NT"], lle["trans"],
                           lle["eps"])
    f_peak, f_lle = d_peak.to_host(), d_lle.to_host()
    s_peak = bh.dbscan_map(peak["tmax"], n, 0.01, IC, R2, [0, 2], 0, 0.0, 100000, peak["trans"],
                           VALS, 1, 0.01)
    s_lle = bh.lle_map(2, lle["tmax"], lle["NT"], n, 0.01, lle["eps"], IC, R2, [0, 2], 0.0,
                       lle["trans"], VALS)
    np.testing.assert_array_equal(f_peak, s_peak)
    assert _same(f_lle, s_lle)
    # and against the oracle on two rows
    for row in (5, 40):
        want = orc.dbscan_cells(peak["tmax"], n, 0.01, IC, R2, [0, 2], 0, 0.0, 100000,
                                peak["trans"], VALS, 1, 0.01, row * n, n)
        np.testing.assert_array_equal(f_peak[row * n:(row + 1) * n], want)


def test_fused_map_1024_passes_against_oracle(bh, orc):
    """A fused map that needs several passes ends with a short one (an
    eighth of a pass, at least 65 536 cells) whose integration runs beside
    the previous pass's DBSCAN on the side stream.  Passes of 400 000 cells
    over a 1024^2 map: 400 000, 400 000, 183 040 and 65 536 cells.  Rows
    across each pass boundary and in the tail equal the oracle (peak counts
    exactly, LLE within 1e-12)."""
    n = 1024
    peak = dict(tmax=10.0, trans=5.0)
    lle = dict(tmax=20.0, NT=0.5, trans=2.0, eps=1e-8)
    bh.set_perturbation(bh.PERT_AXIS_X)
    bh.set_max_cells_per_launch(400000)
    say("fused 1024^2 passes of 400 000")
    d_peak = bh.DeviceArray((n * n,), np.int32).zero_()
    d_lle = bh.DeviceArray((n * n,), np.float64).zero_()
    bh.lle_peak_map_device(d_peak.ptr, d_lle.ptr, n, 0.01, IC, R2, [0, 2], VALS, peak["tmax"],
                           peak["trans"], 0, 100000, 0.01, lle["tmax"], lle["NT"], lle["trans"],
                           lle["eps"])
    f_peak, f_lle = d_peak.to_host(), d_lle.to_host()
    for row in (0, 390, 781, 959, 960, 1023):
        sl = slice(row * n, (row + 1) * n)
        want = orc.dbscan_cells(peak["tmax"], n, 0.01, IC, R2, [0, 2], 0, 0.0, 100000,
                                peak["trans"], VALS, 1, 0.01, row * n, n)
        np.testing.assert_array_equal(f_peak[sl], want, err_msg=f"row {row}")
        wl = orc.lle_cells(2, lle["tmax"], lle["NT"], n, 0.01, lle["eps"], IC, R2, [0, 2], 0.0,
                           lle["trans"], VALS, orc.PERT_AXIS_X, 0, row * n, n)
        got = f_lle[sl]
        assert np.array_equal(np.isnan(got), np.isnan(wl)), f"row {row}: NaN cells differ"
        fin = ~np.isnan(wl)
        assert np.abs(got[fin] - wl[fin]).max() <= 1e-12, f"row {row}"


def test_fused_rejects_window_outside_trajectory(bh):
    d_peak = bh.DeviceArray((16,), np.int32)
    d_lle = bh.DeviceArray((16,), np.float64)
    with pytest.raises(bh.BifhipError) as e:
        bh.lle_peak_map_device(d_peak.ptr, d_lle.ptr, 4, 0.01, IC, R2, [0, 2], VALS, 100.0, 20.0,
                               0, 1000, 0.01, 50.0, 0.5, 10.0, 1e-8)
    assert e.value.code == bh.E_ARG
