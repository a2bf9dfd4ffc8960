"""GPU parity of the trajectory form of the streaming peak finder.

The map kernels find peaks on the integrated trajectory with wave-uniform
plateau / skip masks and take the rising test d[i] > d[i-1] from the
previous sample's compare.  These cases drive every branch of the machine
on real trajectories and compare peaks, counts and LLE with the oracle:
  * linear model x <- x + h*(a*x), factor 1 + h*a over [-2, 1.05]: exact
    sign alternation (a peak every other sample, each followed by a skip),
    decaying alternation that underflows to signed zeros (an endless
    plateau), zeros after one step, growth to inf, alternating growth to
    +-inf and then NaN; with and without the divergence check;
  * Rossler with a slightly below 0, where orbits spiral into a stable
    focus and end in ulp-level steps: about half of the cells hold samples
    that rise onto a plateau inside the recorded window.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = 0.01
LIN_RANGE = [(-2.0 - 1.0) / H, (1.05 - 1.0) / H]   # a: factor 1 + h*a in [-2, 1.05]
LIN_VALS = [0.0, -0.5, 0.3]
LIN_IC = [1.0, 1.0, 1.0]
IC = [0.1, 0.1, 0.1]
FOCUS = [-0.06, -0.005, 5.0, 6.5]                   # (a, c) box around a stable focus


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("max_value", [0.0, 1e30])
def test_linear_alternation_plateaus_and_nan(orc, bh, max_value):
    n = 1024
    say("linear 1D peaks", n, "maxValue", max_value)
    bh.set_model(bh.MODEL_LINEAR)
    cnt, pk = bh.bifurcation1D_to_buffer(20.0, n, H, LIN_IC, LIN_RANGE, [0], 0, max_value, 5.0,
                                         LIN_VALS, 1)
    cap = pk.shape[1]
    wc, wp, _ = orc.peaks_cells(1, 20.0, n, H, LIN_IC, LIN_RANGE, [0], 0, max_value, 5.0,
                                LIN_VALS, 1, 0, n, cap, model=orc.MODEL_LINEAR)
    np.testing.assert_array_equal(cnt, np.maximum(wc, 0))
    for k in range(n):
        np.testing.assert_array_equal(pk[k, :cnt[k]], wp[k, :cnt[k]], err_msg=str(k))
    # the sweep holds alternating cells (many peaks) and peak-free ones
    assert cnt.max() >= 900 and (cnt == 0).sum() > 100


def test_linear_two_cells_per_lane_rows(orc, bh):
    """2^20 cells: the two-cells-per-lane kernel on the same trajectories."""
    n = 1024
    say("linear dbscan map", n)
    bh.set_model(bh.MODEL_LINEAR)
    r = LIN_RANGE + [-1.0, 1.0]
    m = bh.dbscan_map(20.0, n, H, LIN_IC, r, [0, 1], 0, 0.0, 100000, 5.0, LIN_VALS, 1, 0.01)
    m = m.reshape(n, n)
    for row in (0, 517, n - 1):
        want = orc.dbscan_cells(20.0, n, H, LIN_IC, r, [0, 1], 0, 0.0, 100000, 5.0, LIN_VALS, 1,
                                0.01, row * n, n, model=orc.MODEL_LINEAR)
        np.testing.assert_array_equal(m[row], want, err_msg=f"row {row}")


def test_rossler_stable_focus_peaks(orc, bh):
    n = 1024
    say("rossler focus 1D peaks", n)
    r = [FOCUS[0], FOCUS[1]]
    vals = [0.0, 0.2, 5.7]
    # the oracle's series: some cells rise onto a plateau in the window
    steps = 0
    for cell in range(0, n, 64):
        ok, s = orc.bif_cell(cell, n, r, [0], IC, vals, H, 100000, 50000)
        d = np.diff(s)
        steps += int(((d[:-1] > 0) & (d[1:] == 0)).sum())
    assert steps > 0
    cnt, pk = bh.bifurcation1D_to_buffer(500.0, n, H, IC, r, [0], 0, 0.0, 1000.0, vals, 1)
    cap = pk.shape[1]
    wc, wp, _ = orc.peaks_cells(1, 500.0, n, H, IC, r, [0], 0, 0.0, 1000.0, vals, 1, 0, n, cap)
    np.testing.assert_array_equal(cnt, np.maximum(wc, 0))
    for k in range(n):
        np.testing.assert_array_equal(pk[k, :cnt[k]], wp[k, :cnt[k]], err_msg=str(k))


def test_fused_map_stable_focus(orc, bh):
    """Fused LLE + peak-count kernel on the focus box against the oracle."""
    n = 32
    say("fused focus map", n)
    bh.set_perturbation(bh.PERT_AXIS_X)
    vals = [0.2, 0.2, 5.7]
    d_peak = bh.DeviceArray((n * n,), np.int32).zero_()
    d_lle = bh.DeviceArray((n * n,), np.float64).zero_()
    bh.lle_peak_map_device(d_peak.ptr, d_lle.ptr, n, H, IC, FOCUS, [0, 2], vals, 500.0, 1000.0, 0,
                           100000, 0.01, 1000.0, 0.5, 500.0, 1e-8)
    peak, lle = d_peak.to_host(), d_lle.to_host()
    want = orc.dbscan_cells(500.0, n, H, IC, FOCUS, [0, 2], 0, 0.0, 100000, 1000.0, vals, 1, 0.01,
                            0, n * n)
    np.testing.assert_array_equal(peak, want)
    wl = orc.lle_cells(2, 1000.0, 0.5, n, H, 1e-8, IC, FOCUS, [0, 2], 0.0, 500.0, vals,
                       orc.PERT_AXIS_X, 0, 0, n * n)
    nan = np.isnan(wl)
    np.testing.assert_array_equal(np.isnan(lle), nan)
    assert np.abs(lle[~nan] - wl[~nan]).max() <= 1e-12
