"""GPU parity of the FP32 peak maps (bifhip_set_precision(BIFHIP_PREC_FP32)
for the 1D diagram and the 2D cluster maps) through the C ABI.

Standards:
  * against the oracle's FP32 restatement (float state, the same fused FP32
    step, samples widened to double, the reference's peak finder and
    DBSCAN): bit-exact -- every FP32 operation is a correctly rounded add,
    multiply or fma on both sides; KDE over those peaks: count agreement
    >= 99.5%, as for the FP64 KDE map (glibc expf);
  * against the FP64 map: agreement is measured and printed, and checked
    only as a sanity band (the FP32 path has no reference counterpart).
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


def _fp32(bh):
    bh.set_precision(bh.PREC_FP32)


def test_fp32_bifurcation1d_bit_exact(orc, bh):
    """1D diagram, Rossler c in [2, 7] x 400, transient 1000, record 500."""
    say("fp32 bifurcation1D c in [2, 7] x 400")
    _fp32(bh)
    cnt, pk = bh.bifurcation1D_to_buffer(500.0, 400, 0.01, IC, [2.0, 7.0], [2], 0, 0.0, 1000.0,
                                         [0.2, 0.2, 0.0], 1)
    cap = pk.shape[1]
    wc, wp, _ = orc.peaks_cells(1, 500.0, 400, 0.01, IC, [2.0, 7.0], [2], 0, 0.0, 1000.0,
                                [0.2, 0.2, 0.0], 1, 0, 400, cap, fp32=True)
    np.testing.assert_array_equal(cnt, np.maximum(wc, 0))
    for k in range(400):
        np.testing.assert_array_equal(pk[k, :cnt[k]], wp[k, :cnt[k]], err_msg=str(k))
    # the FP64 diagram of the same cells: counts agree on the period-1 part
    bh.set_precision(bh.PREC_FP64)
    c64, _ = bh.bifurcation1D_to_buffer(500.0, 400, 0.01, IC, [2.0, 7.0], [2], 0, 0.0, 1000.0,
                                        [0.2, 0.2, 0.0], 1)
    say("1D peak-count agreement with FP64", float((cnt == c64).mean()))
    assert np.abs(cnt[:40] - c64[:40]).max() <= 1


@pytest.mark.parametrize("wv,pre,max_value", [(0, 1, 0.0), (1, 1, 0.0), (2, 3, 0.0),
                                              (0, 1, 5.0)])
def test_fp32_dbscan_map_bit_exact(orc, bh, wv, pre, max_value):
    """Batched peak stream (pre 1, no check), the per-sample path (pre 3),
    and the divergence check (maxValue 5 makes some cells escape)."""
    n = 20
    args = (100.0, n, 0.01, IC, R2, [0, 2], wv, max_value, 100000, 50.0, VALS, pre, 0.01)
    say("fp32 dbscan map", n, "wv", wv, "pre", pre, "maxValue", max_value)
    _fp32(bh)
    got = bh.dbscan_map(*args)
    want = orc.dbscan_cells(*args, 0, n * n, fp32=True)
    np.testing.assert_array_equal(got, want)
    if max_value:
        assert (want == 0).any()


def test_fp32_dbscan_overflow_and_chunking(orc, bh):
    """Rows of 8 slots send every cell through the exact recompute path,
    which must integrate in FP32 as well; chunked launches change nothing."""
    n = 12
    args = (100.0, n, 0.01, IC, R2, [0, 2], 0, 0.0, 100000, 100.0, VALS, 1, 0.01)
    say("fp32 dbscan overflow + chunking")
    _fp32(bh)
    want = orc.dbscan_cells(*args, 0, n * n, fp32=True)
    bh.set_peak_capacity(8)
    np.testing.assert_array_equal(bh.dbscan_map(*args), want)
    bh.set_peak_capacity(256)
    bh.set_max_cells_per_launch(37)
    np.testing.assert_array_equal(bh.dbscan_map(*args), want)


def test_fp32_dbscan_map_agreement_with_fp64(orc, bh):
    """A 48^2 map at the reference's peak settings: bit-exact against the
    FP32 oracle on sampled rows; its agreement with the FP64 map is printed
    (the FP32 trajectory samples the orbit at other phases, which moves the
    DBSCAN count wherever peak intervals sit one step apart)."""
    n = 48
    args = (500.0, n, 0.01, IC, R2, [0, 2], 0, 0.0, 100000, 1000.0, VALS, 1, 0.01)
    say("fp32 vs fp64 dbscan map", n)
    _fp32(bh)
    m32 = bh.dbscan_map(*args)
    bh.set_precision(bh.PREC_FP64)
    m64 = bh.dbscan_map(*args)
    for row in (0, n // 2, n - 1):
        want = orc.dbscan_cells(*args, row * n, n, fp32=True)
        np.testing.assert_array_equal(m32[row * n:(row + 1) * n], want, err_msg=f"row {row}")
    agree = float((np.clip(m32, 1, 16) == np.clip(m64, 1, 16)).mean())
    say("peak-count class agreement FP32 vs FP64", agree)
    assert agree > 0.5


def test_fp32_kde_map_agreement(orc, bh):
    """KDE over FP32 peaks: the oracle's FP32 peak rows clustered by the
    oracle's kde.  KDE sums glibc expf, which is not correctly rounded, so
    the standard is the FP64 KDE map's: count agreement >= 99.5%."""
    n, cap = 24, 1024
    say("fp32 kde map", n)
    _fp32(bh)
    m = bh.kde_map(200.0, n, 0.01, IC, R2, [0, 2], 0, 0.0, 1000, 200.0, VALS, 1, 10, -15.0,
                   15.0, 0.05)
    cnt, pk, _ = orc.peaks_cells(2, 200.0, n, 0.01, IC, R2, [0, 2], 0, 0.0, 200.0, VALS, 1, 0,
                                 n * n, cap, fp32=True)
    assert cnt.max() < cap
    want = np.array([orc.kde(pk[k, :max(cnt[k], 0)], 1000, 10, -15.0, 15.0, 0.05)
                     for k in range(n * n)])
    agree = float((m == want).mean())
    say("fp32 kde count agreement with the oracle", agree)
    assert agree >= 0.995


def test_fp32_peak_map_refuses_the_memristive_model(bh):
    say("fp32 memristive refused")
    _fp32(bh)
    bh.set_model(bh.MODEL_MEMRISTIVE)
    vals = [0.5, 0.4, 0.3, 0.2, 0.1, 0.05, 0.02, 0.5]
    with pytest.raises(bh.BifhipError):
        bh.dbscan_map(10.0, 4, 0.01, [0.1] * 5, R2, [0, 2], 0, 0.0, 1000, 10.0, vals, 1, 0.01)
