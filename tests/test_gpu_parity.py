"""GPU parity: the HIP kernels through the C ABI against the CPU oracle.

Standards:
  * integer / index results and the polynomial FP64 path (samples, peaks,
    intervals, DBSCAN counts, Gram-Schmidt): bit-exact;
  * LLE / LS (FP64 path): every operation is IEEE-exact except `log`, whose
    device and glibc versions may differ in the last ulp, so |gpu - oracle|
    <= 1e-12 absolute, and the bit-identical fraction is reported;
  * KDE: glibc's expf is not correctly rounded, so count agreement >= 99.5%;
  * FP32 bulk LLE: judged against the FP64 path (error and sign agreement).
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


def _t(bh, a, dtype):
    return bh.DeviceArray.from_host(np.asarray(a), dtype)


def _close_with_nan_class(got, want, atol, what):
    """NaN cells (escaping trajectories) must coincide; NaN sign and payload
    are not compared (x86 and AMD produce different default NaNs).  Finite
    cells within atol; returns the bit-identical fraction."""
    got, want = np.asarray(got), np.asarray(want)
    gn, wn = np.isnan(got), np.isnan(want)
    np.testing.assert_array_equal(gn, wn, err_msg=f"{what}: NaN cells differ")
    fin = ~gn
    err = np.abs(got[fin] - want[fin])
    ident = float((got[fin] == want[fin]).mean()) if fin.any() else 1.0
    say(what, "finite cells", int(fin.sum()), "NaN cells", int(gn.sum()), "max abs err",
        float(err.max()) if err.size else 0.0, "bit-identical", ident)
    assert err.size == 0 or err.max() <= atol
    return ident


# ------------------------------------------------------------ record pass

@pytest.mark.parametrize("wv,pre", [(0, 1), (1, 1), (2, 3)])
def test_series_bit_exact(orc, bh, dev, wv, pre):
    n, tmax, h, trans = 16, 20.0, 0.01, 10.0
    N = int(tmax / h / pre)
    say("series", n, wv, pre)
    data = bh.DeviceArray((n * n * N,), np.float64).zero_()
    flags = bh.DeviceArray((n * n,), np.int32).zero_()
    bh.series_device(data.ptr, flags.ptr, 2, tmax, n, h, IC, R2, [0, 2], wv, 0.0,
                     trans, VALS, pre)
    bh.synchronize()
    got = data.to_host().reshape(n * n, N)
    for cell in range(0, n * n, 5):
        ok, s = orc.bif_cell(cell, n, R2, [0, 2], IC, VALS, h, int(trans / h), N, pre, wv)
        assert ok == 1
        np.testing.assert_array_equal(got[cell], s, err_msg=f"cell {cell}")
    assert not flags.to_host().any()


# ------------------------------------------------------------ peak finder

def _plateau_walks(rng, nb, n):
    d = np.round(np.cumsum(rng.normal(size=(nb, n)), axis=1) * 1.5) / 1.5
    d[::7, 100:140] = d[::7, 100:101]          # long plateaus
    return d


@pytest.mark.parametrize("inplace", [False, True])
def test_peak_finder_device_fuzz(orc, bh, dev, inplace):
    rng = np.random.default_rng(5)
    nb, n = 384, 3000
    d = _plateau_walks(rng, nb, n)
    cnt0 = np.zeros(nb, dtype=np.int32)
    cnt0[::11] = -1
    say("peak_finder", nb, n, inplace)
    D = _t(bh, d.ravel(), np.float64)
    C = _t(bh, cnt0, np.int32)
    P = D if inplace else bh.DeviceArray(D.shape, np.float64).zero_()
    T = bh.DeviceArray(D.shape, np.float64).zero_()
    bh.peak_finder_device(D.ptr, n, nb, C.ptr, P.ptr, T.ptr, 0.01)
    bh.synchronize()
    cnt, pk, tm = C.to_host(), P.to_host().reshape(nb, n), T.to_host().reshape(nb, n)
    for b in range(nb):
        if cnt0[b] == -1:
            assert cnt[b] == 0
            continue
        Pw, pkw, tw = orc.peak_finder(d[b], 0.01)
        assert cnt[b] == Pw, b
        np.testing.assert_array_equal(pk[b, :Pw], pkw)
        np.testing.assert_array_equal(tm[b, :Pw], tw)


def test_peak_finder_reference_kat_gpu(bh, dev):
    import json, os
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))
    case = gold["peak_finder"][0]
    s = np.array(case["series"], dtype=np.float64)
    D = _t(bh, s, np.float64)
    C = bh.DeviceArray((1,), np.int32).zero_()
    P = bh.DeviceArray(D.shape, np.float64).zero_()
    T = bh.DeviceArray(D.shape, np.float64).zero_()
    bh.peak_finder_device(D.ptr, len(s), 1, C.ptr, P.ptr, T.ptr, 0.01)
    bh.synchronize()
    assert int(C.to_host()[0]) == case["count"]
    assert P.to_host()[:2].tolist() == case["peaks"]
    assert T.to_host()[:2].tolist() == case["intervals"]


# ------------------------------------------------------- bifurcation 1D

def test_bifurcation1d_config1_bit_exact(orc, bh):
    """1D diagram, Rossler c in [2,7] x 1000, transient 1000, record 500."""
    say("bifurcation1D c in [2, 7] x 1000")
    cnt, pk = bh.bifurcation1D_to_buffer(500.0, 1000, 0.01, IC, [2.0, 7.0], [2], 0, 0.0, 1000.0,
                                         [0.2, 0.2, 0.0], 1)
    cap = pk.shape[1]
    wc, wp, _ = orc.peaks_cells(1, 500.0, 1000, 0.01, IC, [2.0, 7.0], [2], 0, 0.0, 1000.0,
                                [0.2, 0.2, 0.0], 1, 0, 1000, cap)
    np.testing.assert_array_equal(cnt, np.maximum(wc, 0))
    for k in range(1000):
        np.testing.assert_array_equal(pk[k, :cnt[k]], wp[k, :cnt[k]], err_msg=str(k))
    assert 80 <= cnt.min() and cnt.max() <= 90


def test_bifurcation1d_full_rows_round(orc, bh):
    """Rows of 8 slots cannot hold the ~85 peaks a cell has: the driver goes
    round again with rows that hold every peak a record can have, and the
    diagram is the same, bit for bit."""
    say("bifurcation1D, 8-slot rows then full rows")
    bh.set_peak_capacity(8)
    cnt, pk = bh.bifurcation1D_to_buffer(500.0, 200, 0.01, IC, [2.0, 7.0], [2], 0, 0.0, 1000.0,
                                         [0.2, 0.2, 0.0], 1)
    wc, wp, _ = orc.peaks_cells(1, 500.0, 200, 0.01, IC, [2.0, 7.0], [2], 0, 0.0, 1000.0,
                                [0.2, 0.2, 0.0], 1, 0, 200, pk.shape[1])
    assert cnt.max() > 8
    np.testing.assert_array_equal(cnt, np.maximum(wc, 0))
    for k in range(200):
        np.testing.assert_array_equal(pk[k, :cnt[k]], wp[k, :cnt[k]], err_msg=str(k))
    # a caller's row stride shorter than a point's peaks is an error, not a cut
    with pytest.raises(bh.BifhipError) as e:
        bh.bifurcation1D_to_buffer(500.0, 20, 0.01, IC, [2.0, 7.0], [2], 0, 0.0, 1000.0,
                                   [0.2, 0.2, 0.0], 1, stride=16)
    assert e.value.code == bh.E_LIMIT


def test_bifurcation1d_long_record_integrates_once(bh):
    """A record with more than 1024 peaks in a point: the binding's first
    call (1024-slot rows) is refused with every count filled, and its retry
    with the stride max(counts) copies the kept diagram: as many peak-kernel
    launches as one call with that stride, and the same result."""
    args = (6000.0, 8, 0.01, IC, [2.0, 7.0], [2], 0, 0.0, 100.0, [0.2, 0.2, 0.0], 1)
    say("bifurcation1D long record")
    bh.set_timing(True)
    bh.reset_kernel_times()
    cnt, pk = bh.bifurcation1D_to_buffer(*args)
    first = bh.kernel_time("bif_peaks")[1]
    assert cnt.max() > 1024 and pk.shape[1] == cnt.max()
    bh.reset_kernel_times()
    cnt2, pk2 = bh.bifurcation1D_to_buffer(*args, stride=int(cnt.max()))
    once = bh.kernel_time("bif_peaks")[1]
    say("peak-kernel launches", first, "vs", once)
    assert first == once > 0
    np.testing.assert_array_equal(cnt, cnt2)
    np.testing.assert_array_equal(pk, pk2)


def test_bifurcation1d_file_matches_reference_format(orc, bh, tmp_path):
    say("bifurcation1D file")
    out = tmp_path / "mat.csv"
    bh.set_output_path(out)
    bh.bifurcation1D(100.0, 40, 0.01, 3, IC, [2.0, 7.0], [2], 0, 0.0, 200.0, [0.2, 0.2, 0.0], 3,
                     1)
    N = int(100.0 / 0.01)
    wc, wp, _ = orc.peaks_cells(1, 100.0, 40, 0.01, IC, [2.0, 7.0], [2], 0, 0.0, 200.0,
                                [0.2, 0.2, 0.0], 1, 0, 40, N)
    lines = []
    for k in range(40):
        pv = orc.value_by_idx(k, 40, 2.0, 7.0, 0)
        for j in range(max(wc[k], 0)):
            lines.append(f"{pv:.12g}, {wp[k, j]:.12g}\n")
    assert out.read_text() == "".join(lines)


# ---------------------------------------------------------------- DBSCAN

def _fuzz_peak_sets(rng, nb, stride):
    amps = np.zeros((nb, stride))
    ivls = np.zeros((nb, stride))
    cnt = np.zeros(nb, dtype=np.int32)
    for b in range(nb):
        P = int(rng.choice([0, 1, 2, 3, 5, 17, 63, 64, 65, 130, 200, 256, 257, 300]))
        centers = rng.normal(size=rng.integers(1, 9)) * 0.05
        a = rng.choice(centers, P) + rng.normal(size=P) * rng.choice([0, 0.002, 0.006])
        iv = rng.integers(1, 4, size=P) * 0.01
        if b % 5 == 0:      # pairs placed on the eps boundary
            a[1::2] = a[::2][: len(a[1::2])] + 0.006
            iv[1::2] = iv[::2][: len(iv[1::2])] + 0.008
        amps[b, :P], ivls[b, :P], cnt[b] = a, iv, P
    cnt[::13] = -1
    return amps, ivls, cnt


@pytest.mark.parametrize("max_peaks", [100000, 120])
def test_dbscan_device_fuzz_bit_exact(orc, bh, dev, max_peaks):
    rng = np.random.default_rng(7)
    nb, stride = 600, 320
    amps, ivls, cnt = _fuzz_peak_sets(rng, nb, stride)
    say("dbscan fuzz", nb, max_peaks)
    out = bh.DeviceArray.full((nb,), -99, np.int32)
    A, I, Cn = _t(bh, amps.ravel(), np.float64), _t(bh, ivls.ravel(), np.float64), _t(
        bh, cnt, np.int32)
    bh.dbscan_device(A.ptr, I.ptr, stride, nb, Cn.ptr, max_peaks, 0.01,
                     out.ptr)
    bh.synchronize()
    got = out.to_host()
    for b in range(nb):
        want = 0 if cnt[b] == -1 else orc.dbscan(amps[b, :cnt[b]], ivls[b, :cnt[b]], 0.01,
                                                 max_peaks)
        assert got[b] == want, (b, cnt[b])


def test_dbscan_device_sparse_sets_bit_exact(orc, bh, dev):
    """Chaotic-like peak sets: most peaks isolated, a few small clumps and
    repeats, every peak count from 2 to 256 (one to four 64-peak words)."""
    rng = np.random.default_rng(23)
    sizes = np.arange(2, 257)
    nb, stride = len(sizes), 256
    amps = np.zeros((nb, stride))
    ivls = np.zeros((nb, stride))
    for b, P in enumerate(sizes):
        a = rng.random(P) * rng.choice([0.1, 0.3, 1.0])
        iv = rng.random(P) * 0.05
        k = rng.integers(0, P, size=max(1, P // 8))
        a[k] = a[rng.integers(0, P, size=len(k))] + rng.normal(size=len(k)) * 0.003
        if b % 3 == 0:
            a[-1] = a[0]                     # the last peak is never absorbed
            iv[-1] = iv[0]
        amps[b, :P], ivls[b, :P] = a, iv
    cnt = sizes.astype(np.int32)
    say("dbscan sparse sets", nb)
    out = bh.DeviceArray.full((nb,), -99, np.int32)
    A, I, Cn = _t(bh, amps.ravel(), np.float64), _t(bh, ivls.ravel(), np.float64), _t(
        bh, cnt, np.int32)
    bh.dbscan_device(A.ptr, I.ptr, stride, nb, Cn.ptr, 100000, 0.01, out.ptr)
    bh.synchronize()
    got = out.to_host()
    for b, P in enumerate(sizes):
        assert got[b] == orc.dbscan(amps[b, :P], ivls[b, :P], 0.01), (b, P)


@pytest.mark.parametrize("eps", [0.0, 0.003, 0.01, 0.05, float("inf")])
def test_dbscan_device_chaotic_sets_eps_and_nonfinite(orc, bh, dev, eps):
    """Chaotic-like sets whose isolated peaks the wave kernel finds with its
    amplitude prefilter: for every eps (0: no pair is near; inf: every
    finite pair is), amplitudes on the prefilter bound sqrt(T)(1 + 2^-20)
    and just inside it, exact duplicates, and NaN / infinite amplitudes."""
    rng = np.random.default_rng(31)
    sizes = [2, 3, 40, 63, 64, 65, 100, 128, 129, 190, 256]
    nb, stride = 4 * len(sizes), 256
    amps = np.zeros((nb, stride))
    ivls = np.zeros((nb, stride))
    cnt = np.zeros(nb, dtype=np.int32)
    for b in range(nb):
        P = sizes[b % len(sizes)]
        a = rng.random(P) * 2.0
        iv = rng.integers(1, 6, size=P) * 0.01
        k = rng.integers(0, P, size=max(1, P // 6))
        src = rng.integers(0, P, size=len(k))
        a[k] = a[src] + rng.choice([0.0, 0.00999, 0.0100001, 0.009, 1e-9], size=len(k))
        iv[k] = iv[src]
        if b % 4 == 1 and P > 3:
            a[rng.integers(0, P)] = np.nan
            a[rng.integers(0, P)] = np.inf
        if b % 4 == 2:
            a[-1] = a[0]                     # the last peak is never absorbed
            iv[-1] = iv[0]
        amps[b, :P], ivls[b, :P], cnt[b] = a, iv, P
    say("dbscan chaotic sets eps", eps, nb)
    out = bh.DeviceArray.full((nb,), -99, np.int32)
    A, I, Cn = _t(bh, amps.ravel(), np.float64), _t(bh, ivls.ravel(), np.float64), _t(
        bh, cnt, np.int32)
    bh.dbscan_device(A.ptr, I.ptr, stride, nb, Cn.ptr, 100000, eps, out.ptr)
    bh.synchronize()
    got = out.to_host()
    for b in range(nb):
        P = cnt[b]
        assert got[b] == orc.dbscan(amps[b, :P], ivls[b, :P], eps), (b, P, eps)


def test_wave_kernels_past_one_launch(orc, bh, dev):
    """More cells than one launch of the one-wavefront-per-cell DBSCAN and
    KDE kernels covers (2^24, whose 64 work-items per cell fill 2^30): the
    cells on both sides of the launch boundary and the last ones equal the
    oracle's, so the second launch reads and writes the right rows."""
    n0 = 1 << 24
    nb, stride = n0 + 160, 4
    rng = np.random.default_rng(13)
    cnt = np.zeros(nb, np.int32)
    amps = np.zeros((nb, stride))
    ivls = np.zeros((nb, stride))
    idx = np.r_[np.arange(n0 - 96, nb), [0, 1, 777]]
    for c in idx:
        P = int(rng.integers(-1, stride + 1))
        cnt[c] = P
        if P > 0:
            base = rng.choice([0.5, 2.0, 6.0], size=P)
            amps[c, :P] = base + rng.choice([0.0, 0.004, 0.02], size=P)
            ivls[c, :P] = rng.integers(1, 4, size=P) * 0.01
    say("wave kernels past one launch", nb)
    A, I, Cn = _t(bh, amps.ravel(), np.float64), _t(bh, ivls.ravel(), np.float64), _t(
        bh, cnt, np.int32)
    out = bh.DeviceArray.full((nb,), -99, np.int32)
    bh.dbscan_device(A.ptr, I.ptr, stride, nb, Cn.ptr, 100000, 0.01, out.ptr)
    bh.synchronize()
    got = out.to_host()
    for c in idx:
        P = int(cnt[c])
        want = 0 if P <= 0 else orc.dbscan(amps[c, :P], ivls[c, :P], 0.01)
        assert got[c] == want, ("dbscan", int(c), P)
    assert not (got == -99).any()
    out2 = bh.DeviceArray.full((nb,), -99, np.int32)
    bh.kde_device(A.ptr, stride, nb, Cn.ptr, 100, 10, 0.0, 8.0, 0.05, out2.ptr)
    bh.synchronize()
    got2 = out2.to_host()
    want2 = np.array([0 if cnt[c] <= 0 else orc.kde(amps[c, :cnt[c]], 100, 10, 0.0, 8.0, 0.05)
                      for c in idx])
    # expf may differ from glibc's in the last ulp: the same agreement bar as
    # the KDE fuzz (a misplaced second launch misses most of its cells)
    assert (got2[idx] == want2).mean() >= 0.995
    assert not (got2 == -99).any()


def test_dbscan_device_large_peak_sets(orc, bh, dev):
    """Cells beyond the wave kernel's 256 peaks: up to 16384 peaks take the
    workgroup kernel (peaks staged in LDS up to 3072, read from the rows
    above), larger sets the lane kernel.  Spread sets (mostly singleton
    clusters), few-level sets and near-fixed-point sets (one tight cluster,
    where both kernels stop scanning once every absorbable peak is
    labelled)."""
    rng = np.random.default_rng(11)
    sizes = [257, 900, 3072, 3073, 5000, 6000, 16384, 16390]
    stride = 16400
    nb = len(sizes)
    amps = np.zeros((nb, stride))
    ivls = np.zeros((nb, stride))
    cnt = np.array(sizes, dtype=np.int32)
    for b, P in enumerate(sizes):
        if P > 10000 or P == 6000:
            a = 0.3 + rng.normal(size=P) * 0.0005
            iv = np.full(P, 0.05)
            if P == 16390:
                a[-1] += 1.0        # the last peak, never absorbed, apart
        elif b % 2:
            a = rng.choice([0.0, 0.02, 0.05], P) + rng.normal(size=P) * 0.002
            iv = rng.integers(1, 3, size=P) * 0.01
        else:
            a = rng.random(P) * 0.5
            iv = rng.random(P) * 0.2
        amps[b, :P], ivls[b, :P] = a, iv
    say("dbscan large peak sets", sizes)
    out = bh.DeviceArray.full((nb,), -99, np.int32)
    A, I, Cn = _t(bh, amps.ravel(), np.float64), _t(bh, ivls.ravel(), np.float64), _t(
        bh, cnt, np.int32)
    bh.dbscan_device(A.ptr, I.ptr, stride, nb, Cn.ptr, 1 << 30, 0.01, out.ptr)
    bh.synchronize()
    got = out.to_host()
    for b, P in enumerate(sizes):
        want = orc.dbscan(amps[b, :P], ivls[b, :P], 0.01)
        assert got[b] == want, (b, P, got[b], want)


def test_dbscan_reference_kats_gpu(bh, dev):
    import json, os
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))
    cases = gold["dbscan_eps001_intervals0"]
    stride = 16
    amps = np.zeros((len(cases), stride))
    cnt = np.zeros(len(cases), dtype=np.int32)
    for i, c in enumerate(cases):
        amps[i, :len(c["amps"])] = c["amps"]
        cnt[i] = len(c["amps"])
    A = _t(bh, amps.ravel(), np.float64)
    I = bh.DeviceArray(A.shape, np.float64).zero_()
    Cn = _t(bh, cnt, np.int32)
    out = bh.DeviceArray((len(cases),), np.int32).zero_()
    bh.dbscan_device(A.ptr, I.ptr, stride, len(cases), Cn.ptr, 1 << 30,
                     0.01, out.ptr)
    bh.synchronize()
    assert out.to_host().tolist() == [c["result"] for c in cases]


@pytest.mark.parametrize("max_peaks", [50000, 32])
def test_dbscan_map_config2_rows_bit_exact(orc, bh, max_peaks):
    """Full 512x512 DBSCAN map on the GPU; 6 whole rows (3072 cells spread
    over c) against the oracle."""
    say("dbscan map 512^2", max_peaks)
    m = bh.dbscan_map(500.0, 512, 0.01, IC, R2, [0, 2], 0, 0.0, max_peaks, 1000.0, VALS, 1, 0.01)
    m = m.reshape(512, 512)
    for row in (0, 101, 230, 311, 420, 511):
        want = orc.dbscan_cells(500.0, 512, 0.01, IC, R2, [0, 2], 0, 0.0, max_peaks, 1000.0,
                                VALS, 1, 0.01, row * 512, 512)
        np.testing.assert_array_equal(m[row], want, err_msg=f"row {row}")
    if max_peaks == 32:
        assert (m == 0).mean() > 0.99
    else:
        assert len(np.unique(m)) > 10


def test_dbscan_map_config2_cluster_clamp(orc, bh):
    """Config 2's "capped at 32": bifhip_set_cluster_clamp(32) on the 512x512
    map with every peak admitted gives min(oracle count, 32) on whole rows,
    and the clamp binds somewhere (chaotic cells have 70-85 clusters)."""
    say("dbscan map 512^2, cluster clamp 32")
    bh.set_cluster_clamp(32)
    m = bh.dbscan_map(500.0, 512, 0.01, IC, R2, [0, 2], 0, 0.0, 50000, 1000.0, VALS, 1, 0.01)
    m = m.reshape(512, 512)
    bound = 0
    for row in (5, 260, 480):
        want = orc.dbscan_cells(500.0, 512, 0.01, IC, R2, [0, 2], 0, 0.0, 50000, 1000.0,
                                VALS, 1, 0.01, row * 512, 512)
        np.testing.assert_array_equal(m[row], np.minimum(want, 32), err_msg=f"row {row}")
        bound += int((want > 32).sum())
    assert bound > 0 and m.max() == 32


def test_dbscan_map_two_cells_per_lane(orc, bh):
    """A launch of at least 2^20 cells integrates two cells per lane; a
    1024 x 1024 map (short window) against the oracle on whole rows,
    including the last cell of the launch."""
    say("dbscan map 1024^2, two cells per lane")
    m = bh.dbscan_map(30.0, 1024, 0.01, IC, R2, [0, 2], 0, 0.0, 100000, 20.0, VALS, 1, 0.01)
    m = m.reshape(1024, 1024)
    for row in (0, 333, 1023):
        want = orc.dbscan_cells(30.0, 1024, 0.01, IC, R2, [0, 2], 0, 0.0, 100000, 20.0, VALS, 1,
                                0.01, row * 1024, 1024)
        np.testing.assert_array_equal(m[row], want, err_msg=f"row {row}")


def test_kernel_timers_count_cells(bh):
    """bifhip_kernel_time / bifhip_kernel_cells: a 1024^2 map's peak pass is
    one two-cell launch ("bif_peaks2") of 2^20 cells, a 64^2 map's one
    single-chain launch ("bif_peaks") of 4096 cells; every cell is
    clustered once ("dbscan")."""
    bh.set_timing(True)
    for n, fam in ((1024, "bif_peaks2"), (64, "bif_peaks")):
        bh.reset_kernel_times()
        bh.dbscan_map(10.0, n, 0.01, IC, R2, [0, 2], 0, 0.0, 100000, 5.0, VALS, 1, 0.01)
        ms, launches = bh.kernel_time(fam)
        say("timer", fam, n, ms, launches, bh.kernel_cells(fam))
        assert launches == 1 and ms > 0 and bh.kernel_cells(fam) == n * n
        assert bh.kernel_cells("dbscan") == n * n
        other = "bif_peaks" if fam == "bif_peaks2" else "bif_peaks2"
        assert bh.kernel_time(other)[1] == 0


def test_dbscan_overflow_recompute(orc, bh):
    """Rows of 8 slots force every cell through the exact recompute path."""
    say("dbscan overflow recompute")
    bh.set_peak_capacity(8)
    m = bh.dbscan_map(100.0, 12, 0.01, IC, R2, [0, 2], 0, 0.0, 100000, 100.0, VALS, 1, 0.01)
    want = orc.dbscan_cells(100.0, 12, 0.01, IC, R2, [0, 2], 0, 0.0, 100000, 100.0, VALS, 1,
                            0.01, 0, 144)
    np.testing.assert_array_equal(m, want)


def test_dbscan_divergence_and_chunking(orc, bh):
    say("dbscan divergence + chunking")
    args = (100.0, 20, 0.01, IC, [0.05, 0.45, 2, 10], [0, 2], 0, 5.0, 100000, 50.0, VALS, 1, 0.01)
    full = bh.dbscan_map(*args)
    bh.set_max_cells_per_launch(37)
    chunked = bh.dbscan_map(*args)
    want = orc.dbscan_cells(*args, 0, 400)
    np.testing.assert_array_equal(full, want)
    np.testing.assert_array_equal(chunked, want)
    assert (want == 0).any()


# ------------------------------------------------------------------- KDE

def test_kde_device_agreement(orc, bh, dev):
    rng = np.random.default_rng(9)
    nb, stride = 400, 128
    amps = np.zeros((nb, stride))
    cnt = rng.integers(0, 120, size=nb).astype(np.int32)
    for b in range(nb):
        centers = rng.uniform(0, 10, size=rng.integers(1, 6))
        amps[b, :cnt[b]] = rng.choice(centers, cnt[b]) + rng.normal(size=cnt[b]) * 0.05
    cnt[::17] = -1
    say("kde fuzz", nb)
    out = bh.DeviceArray((nb,), np.int32).zero_()
    A, Cn = _t(bh, amps.ravel(), np.float64), _t(bh, cnt, np.int32)
    bh.kde_device(A.ptr, stride, nb, Cn.ptr, 100, 10, 0.0, 10.0, 0.05,
                  out.ptr)
    bh.synchronize()
    got = out.to_host()
    want = np.array([0 if cnt[b] == -1 else orc.kde(amps[b, :cnt[b]], 100, 10, 0.0, 10.0, 0.05)
                     for b in range(nb)])
    agree = (got == want).mean()
    say("kde agreement", agree)
    assert agree >= 0.995


def test_kde_device_plateaus_zero_runs_and_special_values(orc, bh, dev):
    """Sets that drive the KDE machine's rarer rules, against the oracle:
    sample grids aligned with symmetric peak pairs (sums exactly equal at
    neighbouring points: plateaus that open and close across 64-point
    batches), samples far from every peak (runs of exact-zero sums), NaN
    and infinite peaks, H = 0 and H < 0, a reversed interval, one point per
    peak (kdeSampling 1), and cells with more than 64 peaks."""
    rng = np.random.default_rng(17)
    cases = []
    for t in range(360):
        P = int(rng.integers(3, 150))
        kind = t % 6
        if kind == 0:
            a = (rng.integers(0, 7, size=P) * 0.25).astype(np.float64)
            cases.append((a, 2, -1.0, -1.0 + (2.0 * P - 1) * 0.125, 0.25))
        elif kind == 1:
            a = np.where(np.arange(P) % 2 == 1, 1.0, 3.0)
            cases.append((a, int(rng.choice([1, 2, 3])), 0.0, 4.0, 0.5))
        elif kind == 2:
            a = rng.random(P) * 12
            cases.append((a, 10, 100.0, 130.0, 0.05))
        elif kind == 3:
            a = rng.random(P) * 12
            a[int(rng.integers(0, P))] = [np.nan, np.inf, -np.inf][t % 3]
            cases.append((a, 10, -15.0, 15.0, float(rng.choice([0.05, -0.05]))))
        elif kind == 4:
            a = (rng.integers(0, 5, size=P) * 2.0)
            cases.append((a, 10, 15.0, -15.0, float(rng.choice([0.0, 0.05, 0.2]))))
        else:
            a = rng.integers(0, 3, size=P) * 3.0 + (rng.random(P) - 0.5) * 1e-3
            cases.append((a, int(rng.choice([1, 10, 20])), -15.0, 15.0, 0.05))
    say("kde special sets", len(cases))
    stride = 160
    got = []
    for a, ks, i1, i2, hh in cases:
        amps = np.zeros(stride)
        amps[:len(a)] = a
        A, Cn = _t(bh, amps, np.float64), _t(bh, np.array([len(a)], np.int32), np.int32)
        out = bh.DeviceArray((1,), np.int32).zero_()
        bh.kde_device(A.ptr, stride, 1, Cn.ptr, 1000, ks, i1, i2, hh, out.ptr)
        bh.synchronize()
        got.append(int(out.to_host()[0]))
    want = [orc.kde(a, 1000, ks, i1, i2, hh) for a, ks, i1, i2, hh in cases]
    agree = float(np.mean(np.array(got) == np.array(want)))
    say("kde special agreement", agree)
    assert agree >= 0.995


def test_kde_device_many_peak_chunks(orc, bh, dev):
    """Peak sets of many 64-peak chunks (the kernel loads a cell's peaks 64
    at a time for every batch of sample points): sizes on both sides of
    chunk boundaries up to 1 000 peaks, clustered and spread amplitudes, a
    NaN in a late chunk; counts against the oracle."""
    rng = np.random.default_rng(23)
    sizes = [63, 64, 65, 127, 128, 129, 255, 256, 257, 640, 1000]
    stride = 1024
    got, want = [], []
    for k, P in enumerate(sizes):
        if k % 2:
            a = rng.integers(0, 6, size=P) * 1.5 + rng.normal(size=P) * 0.01
        else:
            a = rng.random(P) * 12
        if P == 640:
            a[600] = np.nan
        amps = np.zeros(stride)
        amps[:P] = a
        A, Cn = _t(bh, amps, np.float64), _t(bh, np.array([P], np.int32), np.int32)
        out = bh.DeviceArray((1,), np.int32).zero_()
        bh.kde_device(A.ptr, stride, 1, Cn.ptr, 100000, 2, -15.0, 15.0, 0.05, out.ptr)
        bh.synchronize()
        got.append(int(out.to_host()[0]))
        want.append(orc.kde(a, 100000, 2, -15.0, 15.0, 0.05))
    say("kde many chunks", got, want)
    assert float(np.mean(np.array(got) == np.array(want))) >= 0.9
    assert sum(g != w for g, w in zip(got, want)) <= 1


def test_kde_map_agreement(orc, bh):
    say("kde map 24^2")
    m = bh.kde_map(200.0, 24, 0.01, IC, R2, [0, 2], 0, 0.0, 1000, 200.0, VALS, 1, 10, -15.0,
                   15.0, 0.05)
    want = orc.kde_cells(200.0, 24, 0.01, IC, R2, [0, 2], 0, 0.0, 1000, 200.0, VALS, 1, 10,
                         -15.0, 15.0, 0.05, 0, 576)
    assert (m == want).mean() >= 0.995
    # the device-output entry point writes the same map
    d = bh.DeviceArray((576,), np.int32).zero_()
    bh.kde_map_device(d.ptr, 200.0, 24, 0.01, IC, R2, [0, 2], 0, 0.0, 1000, 200.0, VALS, 1, 10,
                      -15.0, 15.0, 0.05)
    bh.synchronize()
    np.testing.assert_array_equal(d.to_host(), m)


# --------------------------------------------------------- Gram-Schmidt

def test_gram_schmidt_device_bit_exact(orc, bh, dev):
    rng = np.random.default_rng(11)
    for n in (2, 3, 5, 8):
        cnt = 200
        a = rng.normal(size=(cnt, n, n))
        a[::9, 1] = 0.0
        say("gram_schmidt", n)
        A = _t(bh, a.ravel(), np.float64)
        Bm = bh.DeviceArray(A.shape, np.float64).zero_()
        Dn = bh.DeviceArray((cnt * n,), np.float64).zero_()
        bh.gram_schmidt_device(A.ptr, Bm.ptr, Dn.ptr, n, cnt)
        bh.synchronize()
        b, d = Bm.to_host().reshape(cnt, n, n), Dn.to_host().reshape(cnt, n)
        for i in range(cnt):
            wb, wd = orc.gram_schmidt(a[i])
            np.testing.assert_array_equal(b[i], wb)
            np.testing.assert_array_equal(d[i], wd)


# ------------------------------------------------------------- Lyapunov

@pytest.mark.parametrize("dir_chunk", [-1, 50])
def test_lle2d_fp64_matches_oracle(orc, bh, dir_chunk):
    say("LLE2D fp64", dir_chunk)
    n = 16
    bh.set_direction_chunk(dir_chunk)
    got = bh.lle_map(2, 100.0, 0.5, n, 0.01, 1e-8, IC, R2, [0, 2], 0.0, 50.0, VALS)
    want = orc.lle_cells(2, 100.0, 0.5, n, 0.01, 1e-8, IC, R2, [0, 2], 0.0, 50.0, VALS,
                         orc.PERT_REFERENCE, 200000 if dir_chunk < 0 else dir_chunk, 0, n * n)
    _close_with_nan_class(got, want, 1e-12, "LLE2D")


def test_lle_fp64_divergence_check_and_axis(orc, bh):
    say("LLE1D fp64 maxValue + axis")
    bh.set_perturbation(bh.PERT_AXIS_X)
    got = bh.lle_map(1, 60.0, 0.5, 24, 0.01, 1e-8, IC, [2.0, 12.0], [2], 8.0, 20.0, VALS)
    want = orc.lle_cells(1, 60.0, 0.5, 24, 0.01, 1e-8, IC, [2.0, 12.0], [2], 8.0, 20.0, VALS,
                         orc.PERT_AXIS_X, 0, 0, 24)
    _close_with_nan_class(got, want, 1e-12, "LLE1D maxValue")


def test_ls2d_fp64_matches_oracle(orc, bh):
    say("LS2D fp64")
    n = 8
    got = bh.ls_map(2, 50.0, 0.5, n, 0.01, 1e-6, IC, R2, [0, 2], 0.0, 20.0, VALS)
    want = orc.ls_cells(2, 50.0, 0.5, n, 0.01, 1e-6, IC, R2, [0, 2], 0.0, 20.0, VALS,
                        orc.PERT_REFERENCE, 0, 0, n * n)
    _close_with_nan_class(got, want, 1e-12, "LS2D")


def test_reference_directions_from_host_threads(orc, bh):
    """Maps large enough that the host builds the reference perturbation
    directions (glibc sinf) on several threads: LLE2D on 160^2 cells, LS2D
    on 136^2 and the fused LLE + peak-count map on 160^2 equal the oracle on
    rows spread over the threads' ranges (LLE / LS within 1e-12, peak
    counts exactly)."""
    n = 160
    say("LLE2D reference directions", n)
    got = bh.lle_map(2, 20.0, 0.5, n, 0.01, 1e-8, IC, R2, [0, 2], 0.0, 5.0, VALS).reshape(n, n)
    for r in (0, 53, 107, 159):
        want = orc.lle_cells(2, 20.0, 0.5, n, 0.01, 1e-8, IC, R2, [0, 2], 0.0, 5.0, VALS,
                             orc.PERT_REFERENCE, 200000, r * n, n)
        _close_with_nan_class(got[r], want, 1e-12, f"LLE2D row {r}")
    m = 136
    say("LS2D reference directions", m)
    ls = bh.ls_map(2, 10.0, 0.5, m, 0.01, 1e-6, IC, R2, [0, 2], 0.0, 5.0, VALS)
    for r in (0, 67, 68, 135):
        want = orc.ls_cells(2, 10.0, 0.5, m, 0.01, 1e-6, IC, R2, [0, 2], 0.0, 5.0, VALS,
                            orc.PERT_REFERENCE, 0, r * m, m)
        _close_with_nan_class(ls[r * m:(r + 1) * m].ravel(), np.asarray(want).ravel(), 1e-12,
                              f"LS2D row {r}")
    say("fused map reference directions", n)
    d_peak = bh.DeviceArray((n * n,), np.int32).zero_()
    d_lle = bh.DeviceArray((n * n,), np.float64).zero_()
    bh.lle_peak_map_device(d_peak.ptr, d_lle.ptr, n, 0.01, IC, R2, [0, 2], VALS, 10.0, 5.0, 0,
                           100000, 0.01, 20.0, 0.5, 2.0, 1e-8)
    f_peak, f_lle = d_peak.to_host().reshape(n, n), d_lle.to_host().reshape(n, n)
    for r in (0, 80, 159):
        want = orc.lle_cells(2, 20.0, 0.5, n, 0.01, 1e-8, IC, R2, [0, 2], 0.0, 2.0, VALS,
                             orc.PERT_REFERENCE, 200000, r * n, n)
        _close_with_nan_class(f_lle[r], want, 1e-12, f"fused LLE row {r}")
        wp = orc.dbscan_cells(10.0, n, 0.01, IC, R2, [0, 2], 0, 0.0, 100000, 5.0, VALS, 1, 0.01,
                              r * n, n)
        np.testing.assert_array_equal(f_peak[r], wp, err_msg=f"fused peaks row {r}")


def test_linear_model_closed_forms_gpu(bh):
    import math
    say("linear model closed forms")
    bh.set_model(bh.MODEL_LINEAR)
    bh.set_perturbation(bh.PERT_AXIS_X)
    lam = [0.3, -0.5, -1.0]
    h = 0.01
    lle = bh.lle_map(1, 200.0, 0.5, 4, h, 1e-8, [0, 0, 0], [0.3, 0.3], [0], 0.0, 20.0, lam)
    np.testing.assert_allclose(lle, math.log(1 + h * 0.3) / h, rtol=1e-9)
    ls = bh.ls_map(1, 200.0, 0.5, 3, h, 1e-6, [0, 0, 0], [0.3, 0.3], [0], 0.0, 20.0, lam)
    for row in ls:
        np.testing.assert_allclose(row, [math.log(1 + h * v) / h for v in lam], rtol=1e-9)
    bh.set_precision(bh.PREC_FP32)
    lle32 = bh.lle_map(1, 200.0, 0.5, 4, h, 1e-8, [0, 0, 0], [0.3, 0.3], [0], 0.0, 20.0, lam)
    np.testing.assert_allclose(lle32, math.log(1 + h * 0.3) / h, rtol=1e-5)


def test_lle_fp32_bulk_vs_fp64(bh):
    """FP32 bulk integration with FP64 renormalisation against the FP64
    path on a 48x48 map: regime (chaotic LLE > 0.01) agreement and error."""
    say("LLE fp32 vs fp64 48^2")
    n = 48
    bh.set_perturbation(bh.PERT_AXIS_X)
    f64 = bh.lle_map(2, 300.0, 0.5, n, 0.01, 1e-8, IC, R2, [0, 2], 0.0, 100.0, VALS)
    bh.set_precision(bh.PREC_FP32)
    f32 = bh.lle_map(2, 300.0, 0.5, n, 0.01, 1e-8, IC, R2, [0, 2], 0.0, 100.0, VALS)
    ok = np.isfinite(f64)
    agree = ((f64[ok] > 0.01) == (f32[ok] > 0.01)).mean()
    err = np.abs(f64[ok] - f32[ok])
    say("fp32 regime agreement", agree, "median err", np.median(err), "max", err.max())
    assert agree >= 0.95
    assert np.median(err) < 5e-3


def test_lle_config3_fp32_bulk_vs_fp64(orc, bh):
    """The configured LLE workload: 1024^2 lattice over the (a, c) box,
    transient 500, measure 1000 time units, renormalisation every 0.5,
    perturbation 1e-8 along x.  The FP64 map's rows against the oracle on 4
    sampled rows (|error| <= 1e-12, NaN cells as a class); the FP32 bulk map
    against the FP64 map on every cell.  Bounds: cells NaN in one precision
    only <= 0.2%, regime (LLE > 0.01, NaN not chaotic) agreement >= 98.5%,
    sign agreement >= 98% and mean |error| <= 6e-3 over cells finite in
    both."""
    n = 1024
    bh.set_perturbation(bh.PERT_AXIS_X)
    args = (2, 1000.0, 0.5, n, 0.01, 1e-8, IC, R2, [0, 2], 0.0, 500.0, VALS)
    say("config-3 LLE map fp64", n)
    f64 = bh.lle_map(*args).reshape(n, n)
    rows = (0, 300, 611, 1023)
    for r in rows:
        want = orc.lle_cells(2, 1000.0, 0.5, n, 0.01, 1e-8, IC, R2, [0, 2], 0.0, 500.0, VALS,
                             orc.PERT_AXIS_X, 0, r * n, n)
        _close_with_nan_class(f64[r], want, 1e-12, f"config-3 fp64 row {r}")
    bh.set_precision(bh.PREC_FP32)
    say("config-3 LLE map fp32 bulk", n)
    f32 = bh.lle_map(*args).reshape(n, n)
    a, b = f64.ravel(), f32.ravel()
    # escaping orbits (NaN) near the escape boundary can escape in one
    # precision and not the other; a NaN cell counts as not chaotic
    nan_diff = float((np.isnan(a) != np.isnan(b)).mean())
    ok = np.isfinite(a) & np.isfinite(b)
    err = np.abs(a[ok] - b[ok])
    with np.errstate(invalid="ignore"):
        regime = float(((a > 0.01) == (b > 0.01)).mean())
    sign = float((np.sign(a[ok]) == np.sign(b[ok])).mean())
    say("config-3 fp32 vs fp64: cells", a.size, "NaN in one precision only", nan_diff,
        "regime agreement", regime, "sign agreement (finite)", sign, "mean abs err",
        float(err.mean()), "p99", float(np.quantile(err, 0.99)), "max", float(err.max()))
    assert nan_diff <= 2e-3
    assert regime >= 0.985 and sign >= 0.98 and err.mean() <= 6e-3


# ------------------------------------------------------------- drivers

def test_driver_files_match_reference_format(orc, bh, tmp_path):
    """Every 2D/1D file driver writes the reference's CSV layout with the
    oracle's numbers."""
    from test_abi import _ref_map_text
    say("driver files")
    out = tmp_path / "mat.csv"
    bh.set_output_path(out)
    n = 10
    bh.bifurcation2DDBSCAN(80.0, n, 0.01, 3, IC, R2, [0, 2], 0, 0.0, 100000, 50.0, VALS, 3, 1,
                           0.01)
    want = orc.dbscan_cells(80.0, n, 0.01, IC, R2, [0, 2], 0, 0.0, 100000, 50.0, VALS, 1, 0.01,
                            0, n * n)
    assert out.read_text() == _ref_map_text(n, R2, want, 6)

    bh.LLE2D(40.0, 0.5, n, 0.01, 1e-8, IC, 3, R2, [0, 2], 0, 0.0, 20.0, VALS, 3)
    lle = orc.lle_cells(2, 40.0, 0.5, n, 0.01, 1e-8, IC, R2, [0, 2], 0.0, 20.0, VALS,
                        orc.PERT_REFERENCE, 200000, 0, n * n)
    lines = out.read_text().split("\n")
    assert lines[:2] == ["0.05 0.45", "2 10"]
    body = "".join(lines[2:]).split(", ")
    want_txt = [f"{v:g}" for v in lle]
    assert len(body) == len(want_txt)
    for b_, w_ in zip(body, want_txt):
        # NaN cells print as "nan" or "-nan" depending on the sign bit, which
        # differs between x86 and AMD default NaNs; finite cells to 6 digits
        assert b_ == w_ or (b_ in ("nan", "-nan") and w_ == "nan") or \
            abs(float(b_) - float(w_)) <= 1e-12, (b_, w_)

    bh.LS2D(30.0, 0.5, n, 0.01, 1e-6, IC, 3, R2, [0, 2], 0, 0.0, 10.0, VALS, 3)
    ls = orc.ls_cells(2, 30.0, 0.5, n, 0.01, 1e-6, IC, R2, [0, 2], 0.0, 10.0, VALS,
                      orc.PERT_REFERENCE, 0, 0, n * n)
    for k in range(3):
        p = tmp_path / f"mat.csv{k + 1}.csv"
        lines = p.read_text().split("\n")
        assert lines[:2] == ["0.05 0.45", "2 10"]
        body = [float(v) for v in "".join(lines[2:]).split(", ")]
        np.testing.assert_allclose(body, [float(f"{v:g}") for v in ls[:, k]], rtol=1e-5,
                                   atol=1e-12)

    # LLE1D (Library/hostLibrary.cu:1366-1373): "c, lle" per line at the
    # stream's default 6 significant digits; the directions use the
    # cell's global index (one chunk)
    def same6(b_, w_):
        return b_ == w_ or (b_ in ("nan", "-nan") and w_ == "nan") or \
            abs(float(b_) - float(w_)) <= 1e-12 * max(1.0, abs(float(w_)))

    def c_of(k):
        return orc.value_by_idx(k, n, 2.0, 7.0, 0)

    bh.LLE1D(40.0, 0.5, n, 0.01, 1e-8, IC, 3, [2.0, 7.0], [2], 0, 0.0, 20.0, VALS, 3)
    lle1 = orc.lle_cells(1, 40.0, 0.5, n, 0.01, 1e-8, IC, [2.0, 7.0], [2], 0.0, 20.0, VALS,
                         orc.PERT_REFERENCE, 0, 0, n)
    lines = out.read_text().split("\n")
    assert lines[-1] == "" and len(lines) == n + 1
    for k in range(n):
        c_txt, v_txt = lines[k].split(", ")
        assert c_txt == f"{c_of(k):g}", (k, c_txt)
        assert same6(v_txt, f"{lle1[k]:g}"), (k, v_txt, lle1[k])

    # LS1D (Library/hostLibrary.cu:1898-1908): "c, l1, l2, l3" per line
    bh.LS1D(30.0, 0.5, n, 0.01, 1e-6, IC, 3, [2.0, 7.0], [2], 0, 0.0, 10.0, VALS, 3)
    ls1 = orc.ls_cells(1, 30.0, 0.5, n, 0.01, 1e-6, IC, [2.0, 7.0], [2], 0.0, 10.0, VALS,
                       orc.PERT_REFERENCE, 0, 0, n)
    lines = out.read_text().split("\n")
    assert lines[-1] == "" and len(lines) == n + 1
    for k in range(n):
        f = lines[k].split(", ")
        assert len(f) == 4 and f[0] == f"{c_of(k):g}", (k, lines[k])
        for j in range(3):
            assert same6(f[1 + j], f"{ls1[k, j]:g}"), (k, j, f[1 + j], ls1[k, j])

    # KDE map (Library/hostLibrary.cu:474, 537-545): the header at 12
    # significant digits (a range that needs them), the counts with the 2D
    # separator logic; the counts are the driver's own (its buffer twin) and
    # agree with the oracle (glibc expf is not correctly rounded)
    rk = [0.0512345678901, 0.45, 2.0, 10.0]
    args = (80.0, n, 0.01, 3, IC, rk, [0, 2], 0, 0.0, 1000, 50.0, VALS, 3, 1, 10, -15.0, 15.0,
            0.05)
    bh.bifurcation2DKDE(*args)
    text = out.read_text()
    assert text.split("\n")[:2] == ["0.0512345678901 0.45", "2 10"]
    kde = bh.kde_map(80.0, n, 0.01, IC, rk, [0, 2], 0, 0.0, 1000, 50.0, VALS, 1, 10, -15.0,
                     15.0, 0.05)
    assert text == _ref_map_text(n, rk, kde, 12)
    want = orc.kde_cells(80.0, n, 0.01, IC, rk, [0, 2], 0, 0.0, 1000, 50.0, VALS, 1, 10, -15.0,
                         15.0, 0.05, 0, n * n)
    agree = float((kde == want).mean())
    say("KDE file counts vs oracle agreement", agree)
    assert agree >= 0.99


def test_denis_series_and_file(orc, bh, tmp_path):
    """modelingOneSystemDenis: the interleaved single-system series equals
    the oracle's bit for bit, and the file is "t, x" at 12 digits."""
    say("denis series")
    got = bh.denis_series(30.0, 0.01, 0.2, IC, 1, VALS)
    want = orc.denis(30.0, 0.01, 0.2, IC, 1, VALS)
    np.testing.assert_array_equal(got, want)
    out = tmp_path / "mat.csv"
    bh.set_output_path(out)
    bh.modelingOneSystemDenis(30.0, 0.01, 0.2, 3, IC, 1, VALS, 3)
    lines = out.read_text().splitlines()
    assert len(lines) == len(want)
    assert lines[7] == f"{0.01 * 7:.12g}, {want[7]:.12g}"
