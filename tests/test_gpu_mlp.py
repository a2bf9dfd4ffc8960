"""GPU tests of the surrogate network (BF16 MFMA kernels) and pipeline.

Standards:
  * forward outputs, losses and every gradient tensor of the MFMA kernels
    against the float64 CPU reference (oracle/mlp_ref.py) evaluated with
    the same BF16 rounding points: relative L2 error <= 2e-2 per tensor;
  * training replayed from a captured graph gives bit-identical weights to
    eager launches (deterministic kernels, no atomics on weights);
  * the re-solved cells of the surrogate-refined map equal the plain FP64
    map exactly.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle"))
import mlp_ref as M  # noqa: E402

pytestmark = pytest.mark.gpu

IC = [0.1, 0.1, 0.1]
VALS = [0.2, 0.2, 5.7]
R2 = [0.05, 0.45, 2.0, 10.0]


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _params(seed):
    rng = np.random.default_rng(seed)
    flat = np.zeros(M.N_PARAMS, dtype=np.float32)
    o = 0
    for name, shape in M.SHAPES:
        k = int(np.prod(shape))
        if len(shape) == 2:
            flat[o:o + k] = rng.normal(size=k) * np.sqrt(2.0 / shape[1])
        else:
            flat[o:o + k] = rng.normal(size=k) * 0.05
        o += k
    u = M.unpack(flat)
    u["Wh"][18:] = 0
    u["bh"][18:] = 0
    return M.pack(u).astype(np.float32), (rng.normal(size=64) * 4).astype(np.float32)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


def test_mlp_forward_and_gradients_match_reference(bh):
    assert bh.mlp_param_count() == M.N_PARAMS
    rng = np.random.default_rng(3)
    n_lat, rows = 64, 300
    flat, four = _params(4)
    cells = rng.integers(0, n_lat * n_lat, size=rows)
    cls = rng.integers(0, 16, size=rows)
    lle = rng.normal(size=rows).astype(np.float32)
    lle[::9] = np.nan
    say("mlp gradient check", rows)
    g, loss, out = bh.mlp_gradient(n_lat, flat, four, cells, cls, lle, 0.5, chaos_std=0.2)
    wl, wout, wg = M.forward_backward(flat.astype(np.float64), four, cells, n_lat, cls, lle, 0.5,
                                      bf=True, chaos_std=0.2)
    say("losses gpu", loss, "ref", wl)
    np.testing.assert_allclose(loss, wl, rtol=2e-3)
    assert _rel(out[:, :18], wout[:, :18]) < 1e-2
    o = 0
    for name, shape in M.SHAPES:
        k = int(np.prod(shape))
        e = _rel(g[o:o + k], wg[o:o + k])
        say("grad", name, "rel err", e)
        assert e < 2e-2, name
        o += k


def _synthetic_labels(n):
    """A 2D map with curved class boundaries and a smooth LLE field."""
    j, i = np.divmod(np.arange(n * n), n)
    u, v = i / (n - 1), j / (n - 1)
    cls = (np.floor(4 * (u + 0.3 * np.sin(3 * v))) % 4).astype(np.int32) * 3
    cls[(u - 0.7) ** 2 + (v - 0.3) ** 2 < 0.02] = 15
    lle = (np.sin(4 * u) * np.cos(3 * v)).astype(np.float32)
    return cls, lle


def test_adam_kernel_matches_reference_adam(bh):
    """k_adam (through bifhip_mlp_adam) against the float64 Adam of
    oracle/mlp_ref.py with the kernel's float32 constants: six steps, the
    gradient as three partials, magnitudes from 1e-9 (where eps dominates)
    to 10, a quarter of the entries held at zero.  Tolerance: float32
    rounding of the update (relative 2e-6 of p plus 2e-9 absolute; the first
    moment within 1e-6 of the gradients' scale, the second within a
    relative 1e-5); the BF16
    copy must be the round-to-nearest-even of the FP32 parameter, bit for
    bit."""
    rng = np.random.default_rng(12)
    n, steps, splits = 5000, 6, 3
    f = lambda x: float(np.float32(x))
    lr, b1, b2, eps = f(1e-3), f(0.9), f(0.999), f(1e-8)
    p0 = rng.normal(size=n).astype(np.float32)
    mag = 10.0 ** rng.uniform(-9, 1, size=n)
    grads = (rng.normal(size=(steps, splits, n)) * mag).astype(np.float32)
    keep = (np.arange(n) % 4 != 1).astype(np.uint8)
    say("adam", n, steps, splits)
    p, m, v, wbf = bh.mlp_adam(p0, np.zeros(n), np.zeros(n), grads, keep, lr, b1, b2, eps)
    pr, mr, vr = p0.astype(np.float64), np.zeros(n), np.zeros(n)
    for k in range(steps):
        gk = grads[k].astype(np.float32).sum(axis=0, dtype=np.float32)
        pr, mr, vr = M.adam_step(pr, mr, vr, gk, k + 1, lr, b1, b2, eps, keep=keep)
    np.testing.assert_allclose(p, pr, rtol=2e-6, atol=2e-9)
    # m sums gradients of either sign: its float32 rounding is relative to
    # the gradients' scale, not to m itself
    gscale = np.abs(grads).sum(axis=1).max(axis=0)
    assert np.all(np.abs(m - mr) <= 1e-5 * np.abs(mr) + 1e-6 * gscale)
    np.testing.assert_allclose(v, vr, rtol=1e-5, atol=1e-36)
    assert not p[keep == 0].any()
    got_bf = (wbf.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    np.testing.assert_array_equal(got_bf, M.bf16(p))
    # the update moved every live parameter by up to lr per step
    live = keep == 1
    assert np.all(np.abs(p[live] - p0[live]) <= steps * lr * 1.01)


def test_training_graph_equals_eager(bh):
    n = 32
    cls, lle = _synthetic_labels(n)
    dc = bh.DeviceArray.from_host(cls, np.int32)
    dl = bh.DeviceArray.from_host(lle, np.float32)
    say("train graph vs eager")
    l1 = bh.mlp_train_device(n, dc.ptr, dl.ptr, 3, 256, 1e-3, 1.0, 7, 4.0, graph=True)
    p1, _ = bh.mlp_get_params()
    l2 = bh.mlp_train_device(n, dc.ptr, dl.ptr, 3, 256, 1e-3, 1.0, 7, 4.0, graph=False)
    p2, _ = bh.mlp_get_params()
    np.testing.assert_array_equal(p1, p2)
    say("losses", l1, l2)


def test_training_learns_a_synthetic_map(bh):
    n, nf = 64, 128
    cls, lle = _synthetic_labels(n)
    dc = bh.DeviceArray.from_host(cls, np.int32)
    dl = bh.DeviceArray.from_host(lle, np.float32)
    say("train synthetic 64^2, 20 epochs")
    loss = bh.mlp_train_device(n, dc.ptr, dl.ptr, 20, 256, 1e-3, 1.0, 1, 4.0)
    say("final losses", loss)
    pc = bh.DeviceArray((nf * nf,), np.uint8)
    ch = bh.DeviceArray((nf * nf,), np.uint8)
    conf = bh.DeviceArray((nf * nf,), np.float32)
    pl = bh.DeviceArray((nf * nf,), np.float32)
    bh.mlp_infer_device(nf, 0.0, 1.0, 0.0, pc.ptr, ch.ptr, conf.ptr, pl.ptr)
    want_c, want_l = _synthetic_labels(nf)
    acc = (pc.to_host() == want_c).mean()
    mae = np.abs(pl.to_host() - want_l).mean()
    cacc = (ch.to_host() == (want_l > 0)).mean()
    say("fine-grid accuracy", acc, "LLE mae", mae, "chaos-flag accuracy", cacc)
    assert acc > 0.93 and mae < 0.1 and cacc > 0.93
    assert loss[0] < 0.3


# cap: peak-row capacity (8 sends every re-solved cell through the overflow
# recompute, which gathers cell indices from the device lists); pert: the
# reference perturbation directions read the listed cells' indices back;
# margin: FP32 bulk LLE for regime-only cells this far from the threshold
# (negative: all FP64)
# reading: what a cell's regime is (0 class and chaos bit, 1 chaos bit, 2
# class); quad: the coarse-quad chaos rule's confidence threshold (0 off)
@pytest.mark.parametrize("partial,cap,pert,margin,reading,quad",
                         [(True, 0, 1, 0.02, 0, 0.0), (False, 0, 1, 0.02, 0, 0.0),
                          (True, 8, 1, 0.02, 0, 0.0), (True, 0, 0, 0.02, 0, 0.0),
                          (True, 0, 1, -1.0, 0, 0.0), (True, 0, 1, 0.02, 0, 2.0),
                          (True, 0, 1, 0.01, 1, 2.0), (True, 0, 1, 0.02, 2, 0.0)])
def test_surrogate_pipeline_small(bh, partial, cap, pert, margin, reading, quad):
    nf, nc = 96, 24
    peak = dict(tmax=40.0, trans=20.0)
    lle = dict(tmax=50.0, NT=0.5, trans=10.0, eps=1e-8)
    bh.set_perturbation(pert)
    bh.set_partial_resolve(partial)
    bh.set_lle32_margin(margin)
    bh.set_regime_reading(reading)
    bh.set_quad_confidence(quad)
    if cap:
        bh.set_peak_capacity(cap)
    say("surrogate pipeline", nf, nc, "partial", partial, "cap", cap, "pert", pert, "margin",
        margin, "reading", reading, "quad", quad)
    d_pk = bh.DeviceArray((nf * nf,), np.int32).zero_()
    d_ll = bh.DeviceArray((nf * nf,), np.float64).zero_()
    d_rs = bh.DeviceArray((nf * nf,), np.uint8).zero_()
    st = bh.surrogate_map_device(d_pk.ptr, d_ll.ptr, d_rs.ptr, nf, nc, 0.01, IC, R2, [0, 2],
                                 VALS, peak["tmax"], peak["trans"], 0, 100000, 0.01,
                                 lle["tmax"], lle["NT"], lle["trans"], lle["eps"], epochs=20,
                                 batch=128, seed=3, validation=0.05)
    say("stats", st)
    t_pk = bh.DeviceArray((nf * nf,), np.int32).zero_()
    t_ll = bh.DeviceArray((nf * nf,), np.float64).zero_()
    bh.lle_peak_map_device(t_pk.ptr, t_ll.ptr, nf, 0.01, IC, R2, [0, 2], VALS, peak["tmax"],
                           peak["trans"], 0, 100000, 0.01, lle["tmax"], lle["NT"], lle["trans"],
                           lle["eps"])
    # the FP32 bulk LLE map of the lattice (the surrogate's regime re-solves
    # run the same kernel on listed cells)
    t_32 = bh.DeviceArray((nf * nf,), np.float64).zero_()
    bh.set_precision(bh.PREC_FP32)
    try:
        bh.lle_map_device(t_32.ptr, 2, lle["tmax"], lle["NT"], nf, 0.01, lle["eps"], IC, R2,
                          [0, 2], 0.0, lle["trans"], VALS)
    finally:
        bh.set_precision(bh.PREC_FP64)
    pk, ll, rs = d_pk.to_host(), d_ll.to_host(), d_rs.to_host()
    tpk, tll, t32 = t_pk.to_host(), t_ll.to_host(), t_32.to_host()
    full = (rs == 1) | (rs == 3)
    assert (rs == 1).sum() == st["resolved"]
    assert (rs == 2).sum() == st["peak_only"]
    assert (rs == 3).sum() == st["validation"] > 0
    assert ((rs >= 4) & (rs <= 7)).sum() == st["lle_only"]
    assert ((rs == 5) | (rs == 7)).sum() == st["lle_only_recounted"]
    assert ((rs == 6) | (rs == 7)).sum() == st["lle_only_fp32"]
    if not partial:
        assert st["peak_only"] == 0 and st["lle_only"] == 0
    if margin < 0:
        assert st["lle_only_fp32"] == 0
    elif partial and reading != 2:
        assert st["lle_only_fp32"] > 0
    # cells re-solved for the regime only carry the FP64 LLE (4, 5: FP32
    # LLE within the margin of the threshold, or NaN) or the FP32 bulk LLE
    # (6, 7: outside it), and the exact peak count where that regime differs
    # from the predicted one (5, 7)
    lo = (rs == 4) | (rs == 5)
    same_lo = (np.isnan(ll[lo]) & np.isnan(tll[lo])) | (ll[lo] == tll[lo])
    assert same_lo.all()
    lo32 = (rs == 6) | (rs == 7)
    np.testing.assert_array_equal(ll[lo32], t32[lo32])
    with np.errstate(invalid="ignore"):
        assert (np.abs(t32[lo32] - 0.01) >= margin).all()
        if margin >= 0:
            assert not (np.abs(t32[lo] - 0.01) >= margin).any()
    rc = (rs == 5) | (rs == 7)
    np.testing.assert_array_equal(pk[rc], tpk[rc])
    # cells re-solved in full equal the plain FP64 map exactly; cells
    # re-solved for the class carry the exact peak count
    np.testing.assert_array_equal(pk[full], tpk[full])
    same = (np.isnan(ll[full]) & np.isnan(tll[full])) | (ll[full] == tll[full])
    assert same.all()
    np.testing.assert_array_equal(pk[rs == 2], tpk[rs == 2])
    cls_t = np.clip(tpk, 1, 16)
    cls_s = np.clip(pk, 1, 16)
    with np.errstate(invalid="ignore"):
        ch_ok = (tll > 0.01) == (ll > 0.01)
    reg = (cls_t == cls_s) & ch_ok
    say("regime agreement", reg.mean(), "chaos", ch_ok.mean(), "class",
        (cls_t == cls_s).mean(), "full", full.mean(), "peak only", (rs == 2).mean())
    if reading == 1:
        # chaos bit only: no class re-solves and no recounts
        assert st["resolved"] == 0 and st["peak_only"] == 0
        assert st["lle_only_recounted"] == 0
        assert ch_ok.mean() > 0.9
    elif reading == 2:
        # class only: no LLE re-solves
        assert st["resolved"] == 0 and st["lle_only"] == 0
        assert (cls_t == cls_s).mean() > 0.9
    else:
        assert reg.mean() > 0.9
    if quad > 1.0:
        # every cell of a coarse quad whose exact corners disagree on the
        # chaos bit has its LLE re-solved (the network-independent half of
        # the rule); the coarse map is the plain map of the coarse lattice
        c_pk = bh.DeviceArray((nc * nc,), np.int32).zero_()
        c_ll = bh.DeviceArray((nc * nc,), np.float64).zero_()
        bh.lle_peak_map_device(c_pk.ptr, c_ll.ptr, nc, 0.01, IC, R2, [0, 2], VALS, peak["tmax"],
                               peak["trans"], 0, 100000, 0.01, lle["tmax"], lle["NT"],
                               lle["trans"], lle["eps"])
        with np.errstate(invalid="ignore"):
            q = (c_ll.to_host() > 0.01).reshape(nc, nc).astype(int)
        s4 = q[:-1, :-1] + q[1:, :-1] + q[:-1, 1:] + q[1:, 1:]
        mixed = (s4 > 0) & (s4 < 4)
        k = np.minimum(np.arange(nf) * (nc - 1) // (nf - 1), nc - 2)
        mq = mixed[np.ix_(k, k)].ravel()
        say("cells in mixed coarse quads", int(mq.sum()))
        assert mq.any()
        lle_resolved = (rs == 1) | ((rs >= 4) & (rs <= 7))
        assert lle_resolved[mq].all()


def _regimes(peak, lle):
    with np.errstate(invalid="ignore"):
        return np.clip(peak, 1, 16), lle > 0.01


def test_surrogate_gain_over_plain_map(bh):
    """What the trained network gains, on a scaled-down surrogate map (512^2
    fine lattice, 128^2 coarse map, 20 epochs): it re-solves under 60% of
    the plain map's work (0.38 measured) at >= 98% per-cell regime agreement
    (98.5% measured), while the same pipeline with an untrained network
    (0 epochs) re-solves everything.  Work is counted in plain-map cells: a
    full re-solve 1, a peak-count re-solve 0.6 (also the recount of a
    regime-only cell whose regime changed), an FP32 bulk LLE 0.36 (every
    regime-only cell), an FP64 LLE 0.86 (those near the threshold): their
    kernel times per cell relative to the fused kernel plus DBSCAN.
    At the configured scale (4096^2 from a 256^2 coarse map, bench.py) the
    agreement is >= 99%; a coarser fine lattice has more regime boundaries
    per cell."""
    nf, nc = 512, 128
    peak = dict(tmax=500.0, trans=1000.0)
    lle = dict(tmax=1000.0, NT=0.5, trans=500.0, eps=1e-8)
    bh.set_perturbation(bh.PERT_AXIS_X)
    t_pk = bh.DeviceArray((nf * nf,), np.int32).zero_()
    t_ll = bh.DeviceArray((nf * nf,), np.float64).zero_()
    bh.lle_peak_map_device(t_pk.ptr, t_ll.ptr, nf, 0.01, IC, R2, [0, 2], VALS, peak["tmax"],
                           peak["trans"], 0, 50000, 0.01, lle["tmax"], lle["NT"], lle["trans"],
                           lle["eps"])
    t_cls, t_ch = _regimes(t_pk.to_host(), t_ll.to_host())
    res = {}
    for epochs in (20, 0):
        d_pk = bh.DeviceArray((nf * nf,), np.int32).zero_()
        d_ll = bh.DeviceArray((nf * nf,), np.float64).zero_()
        say("surrogate gain", nf, nc, "epochs", epochs)
        st = bh.surrogate_map_device(d_pk.ptr, d_ll.ptr, None, nf, nc, 0.01, IC, R2, [0, 2],
                                     VALS, peak["tmax"], peak["trans"], 0, 50000, 0.01,
                                     lle["tmax"], lle["NT"], lle["trans"], lle["eps"],
                                     epochs=epochs, batch=512, seed=5, validation=0.0)
        s_cls, s_ch = _regimes(d_pk.to_host(), d_ll.to_host())
        agree = float(((s_cls == t_cls) & (s_ch == t_ch)).mean())
        lle64 = st["lle_only"] - st["lle_only_fp32"]
        work = (st["resolved"] + 0.6 * (st["peak_only"] + st["lle_only_recounted"]) +
                0.36 * st["lle_only"] + 0.86 * lle64) / (nf * nf)
        res[epochs] = (agree, work)
        say("epochs", epochs, "agreement", agree, "work vs plain map", work, st)
    assert res[20][0] >= 0.98 and res[0][0] >= 0.999
    assert res[20][1] < 0.6 and res[0][1] > 0.95
    assert res[20][1] < res[0][1] - 0.3
