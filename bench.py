#!/usr/bin/env python3
"""Headline benchmark: grid points/s of the 2D LLE + peak-count map,
surrogate-refined versus plain FP64, at >= 99% per-cell regime agreement.

Rossler system, b = 0.2, x0 = (0.1, 0.1, 0.1), h = 0.01, (a, c) lattice over
[0.05, 0.45] x [2, 10] (synthetic parameter lattice, nothing is loaded):
  * peak count: DBSCAN cluster count of (amplitude, interval) peak pairs,
    transient 1000, record 500 time units, eps = 0.01, no peak threshold;
  * LLE: Benettin two-trajectory, transient 500, measure 1000 time units,
    renormalisation every 0.5, perturbation 1e-8 along x.
Regime of a cell = (peak-count class clamp(count, 1, 16), LLE > 0.01).

Workloads (one step = one whole map, inputs resident, outputs in HBM):
  surrogate  (default) 4096^2 fine map: plain FP64 coarse map on 256^2, a
             20-epoch BF16-MFMA network, inference, FP64 re-solve of the
             flagged cells.  Agreement and speedup are measured against the
             plain FP64 map of the same 4096^2 lattice (computed once).
  plain      the plain FP64 fused LLE + peak-count map on an n x n lattice.
  lle        the 1024^2 LLE map with FP32 bulk integration, timed against
             the plain FP64 LLE map, with its error and sign agreement.
  dbscan     the 512^2 DBSCAN peak-count map (capped at 32), plain FP64,
             with sampled rows checked bit for bit against the oracle.
  kde        the 512^2 KDE peak-count map, plain FP64, with sampled rows'
             count agreement with the oracle.
  drivers    every file-writing driver of the reference at its configuration
             size (host inputs, CSV written): wall and kernel time per call.
  bif1d      the 1000-point 1D bifurcation diagram through the reference's
             driver, every cell checked against the oracle (CPU baseline).
  tiles      Lyapunov spectrum of 4096 uneven 16x16 refinement tiles (FP32
             variational state, FP64 Gram-Schmidt), autotuned, one
             persistent graph-replayed launch per step.
  copy       a 1 GiB device copy: the reference that scales the HBM byte
             counters (FETCH_SIZE, WRITE_SIZE) of the profiles.
Prints one JSON line.  --gpus above 1 is refused: this build uses one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cudabifurcation3.0_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

IC = [0.1, 0.1, 0.1]
VALS = [0.2, 0.2, 5.7]
RANGES = [0.05, 0.45, 2.0, 10.0]
H = 0.01
PEAK = dict(tMax=500.0, transient=1000.0, eps=0.01)
LLE = dict(tMax=1000.0, NT=0.5, transient=500.0, eps=1e-8)
CHAOS = 0.01
FP64_VALU_PEAK = 78.6e12     # MI355X FP64 vector, FMA counted as 2
FP32_VALU_PEAK = 157.3e12    # MI355X FP32 vector (published), FMA counted as 2
STEP_FLOP = 12               # Rossler sequential Euler step, no contraction
RENORM_FLOP = 20
TILE_STEP_FLOP = 48          # state step + 3 tangent-vector steps
METRIC = "grid points/s, 2D LLE+peak map; surrogate-refined vs plain FP64 @≥99% agree"


def lle_flop_per_cell():
    skip = int(LLE["transient"] / H)
    nNT = int(LLE["NT"] / H)
    nb = int(LLE["tMax"] / LLE["NT"])
    return STEP_FLOP * (skip + 2 * nb * nNT) + RENORM_FLOP * nb


def n_rec():
    return int(PEAK["tMax"] / H)


def plain_map(bh, n, d_peak, d_lle):
    """Plain FP64 map: one fused pass gives both the LLE (LLE2D semantics)
    and the DBSCAN peak count (bifurcation2DDBSCAN semantics)."""
    bh.lle_peak_map_device(d_peak.ptr, d_lle.ptr, n, H, IC, RANGES, [0, 2], VALS, PEAK["tMax"],
                           PEAK["transient"], 0, n_rec(), PEAK["eps"], LLE["tMax"], LLE["NT"],
                           LLE["transient"], LLE["eps"])


def surrogate_map(bh, n, nc, d_peak, d_lle, d_res, seed, conf, validation, batch):
    return bh.surrogate_map_device(d_peak.ptr, d_lle.ptr, d_res.ptr, n, nc, H, IC, RANGES,
                                   [0, 2], VALS, PEAK["tMax"], PEAK["transient"], 0, n_rec(),
                                   PEAK["eps"], LLE["tMax"], LLE["NT"], LLE["transient"],
                                   LLE["eps"], epochs=20, batch=batch, lr=1e-3, seed=seed,
                                   conf=conf, chaos=CHAOS, validation=validation)


def regimes(np, peak, lle):
    """(peak-count class, chaotic) per cell; a NaN LLE (escaping orbit)
    counts as not chaotic."""
    with np.errstate(invalid="ignore"):
        return np.clip(peak, 1, 16), lle > CHAOS


def boundary_fraction(np, cls, ch, n):
    """Cells with a 4-neighbour of another regime in the reference map."""
    c = cls.reshape(n, n)
    h = ch.reshape(n, n)
    b = np.zeros((n, n), dtype=bool)
    for a in (c, h):
        d0 = a[:, 1:] != a[:, :-1]
        d1 = a[1:, :] != a[:-1, :]
        b[:, 1:] |= d0
        b[:, :-1] |= d0
        b[1:, :] |= d1
        b[:-1, :] |= d1
    return float(b.mean())


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n, rows, repeats=3):
    """Oracle (C restatement of the reference's per-trajectory code, built
    with gcc -O3 -ffp-contract=off, OpenMP over cells) on `rows` whole rows of
    the same map: DBSCAN peak count and LLE as the reference computes them
    (two separate trajectories).  Median of `repeats` runs."""
    import orclib as O
    O.lib()
    step = max(1, n // rows)
    sel = list(range(0, n, step))[:rows]
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        for r in sel:
            O.dbscan_cells(PEAK["tMax"], n, H, IC, RANGES, [0, 2], 0, 0.0, n_rec(),
                           PEAK["transient"], VALS, 1, PEAK["eps"], r * n, n)
            O.lle_cells(2, LLE["tMax"], LLE["NT"], n, H, LLE["eps"], IC, RANGES, [0, 2], 0.0,
                        LLE["transient"], VALS, O.PERT_AXIS_X, 0, r * n, n)
        times.append(time.perf_counter() - t0)
    cells = len(sel) * n
    dt = sorted(times)[len(times) // 2]
    return {"value": cells / dt, "unit": "grid points/s", "cores": O.lib().orc_threads(),
            "cpu_model": cpu_model(), "kind": "port", "build": "gcc -O3 -ffp-contract=off "
            "-fopenmp", "runs_s": times,
            "sample": f"{cells} cells ({len(sel)} evenly spaced rows of the {n}x{n} lattice), "
                      f"DBSCAN peak count + LLE per cell, median of {repeats} runs: {dt:.1f} s"}


# HBM bytes per cell of k_lle_peaks at the benchmark's settings, from the
# FETCH_SIZE / WRITE_SIZE counter runs of the plain 1024^2 map scaled by the
# same counters on a 1 GiB copy (FETCH x 2.0, WRITE x 1.0): 478 KiB fetched
# and 1 495 424 KiB written per 1 048 576-cell launch.  Algorithmic bytes are
# about 1 372 per cell (85 peaks x 16 B, plus 12 B of results): the peaks
# reach the rows four at a time from an LDS stage, so the writes are 1.06x.
LLE_PEAKS_HBM_BYTES_PER_CELL = (478 + 1495424) * 1024 / 1048576


def roofline(bh, cells_total, kernel="lle_peaks"):
    """Achieved FP64 FLOP/s of the fused integrator over all its launches:
    algorithmic FLOP (12 per Rossler step, 20 per renormalisation; peak
    detection compares are not counted) / summed HIP-event launch time."""
    ms, n = bh.kernel_time(kernel)
    avg = ms / max(n, 1)
    flop = lle_flop_per_cell() * cells_total
    ach = flop / (ms / 1000) / 1e12 if ms > 0 else 0.0
    return {"bound": "valu_fp64", "kernel": "k_lle_peaks", "achieved": ach,
            "peak": FP64_VALU_PEAK / 1e12, "unit": "TFLOP/s", "frac": ach / (FP64_VALU_PEAK / 1e12),
            "traffic": LLE_PEAKS_HBM_BYTES_PER_CELL * cells_total / max(n, 1),
            "traffic_note": "HBM bytes per launch from the counter profile (per-cell figure x "
                            "cells per launch)",
            "flop_per_launch": flop / max(n, 1), "avg_launch_ms": avg,
            "launches": n, "cells": cells_total,
            "note": "FP64 vector ALU (no MFMA on this path); no-contraction FP64 reaches at "
                    "most 50% of the FMA-counted peak"}


def tiles_fp32_vs_fp64(bh, np, f32, f64):
    """Per-cell error of the FP32 tile spectrum against the FP64 tile
    kernel (FP64 state and tangent vectors, the reference's Gram-Schmidt)."""
    a, b = f32.reshape(-1, 3), f64.reshape(-1, 3)
    fa, fb = np.isfinite(a).all(axis=1), np.isfinite(b).all(axis=1)
    fin = fa & fb
    e = np.abs(a[fin] - b[fin])
    ssum = np.abs(a[fin].sum(axis=1) - b[fin].sum(axis=1)) / \
        np.maximum(np.abs(b[fin].sum(axis=1)), 1e-12)
    return {"cells_compared": int(fin.sum()), "finite_in_one_only": float((fa != fb).mean()),
            "regime_agreement": float(((a[fin, 0] > CHAOS) == (b[fin, 0] > CHAOS)).mean()),
            "abs_err_median": np.median(e, axis=0).tolist(),
            "abs_err_p90": np.quantile(e, 0.9, axis=0).tolist(),
            "abs_err_p99": np.quantile(e, 0.99, axis=0).tolist(),
            "spectrum_sum_rel_err_median": float(np.median(ssum)),
            "spectrum_sum_rel_err_p99": float(np.quantile(ssum, 0.99))}


def run_config4(bh, np, steps=1, warmup=0, keep=None):
    """Config 4: 4096 tiles of 16x16 cells on the 4096^2 lattice, seeded
    lengths U[200, 2000] time units, FP32 variational state, FP64
    Gram-Schmidt every 0.5; autotuned, one persistent graph-replayed launch
    per map; then the same tiles with the FP64 tile kernel for the error.
    `keep` (a dict) receives the tiles and the FP64 spectra."""
    nt, w, n_pts = 4096, 16, 4096
    tiles = bh.tiles_generate(nt, w, w, n_pts, 0, 1, 200.0, 2000.0, LLE["NT"])
    nsteps = int(np.asarray(tiles[2], dtype=np.int64).sum()) * int(LLE["NT"] / H) * w * w
    d = bh.DeviceArray((nt * w * w * 3,), np.float64)
    bh.set_tile_config(0)
    args_ = (d.ptr, tiles, w, w, n_pts, RANGES, [0, 2], IC, VALS, H, LLE["NT"], 0.0)
    bh.ls_tiles_device(*args_, autotune=True)
    tune = bh.tile_autotune_log().tolist()
    for _ in range(warmup):
        bh.ls_tiles_device(*args_)
    bh.synchronize()
    bh.set_timing(True)
    bh.reset_kernel_times()
    t0 = time.perf_counter()
    for _ in range(steps):
        bh.ls_tiles_device(*args_)
    bh.synchronize()
    ms = (time.perf_counter() - t0) * 1000 / steps
    kms, nl = bh.kernel_time("ls_tiles")
    f32 = d.to_host()
    block, cpl, _ = bh.get_tile_config()
    lanes = bh.get_tile_lanes()
    bh.set_tile_precision(bh.PREC_FP64)
    try:
        bh.set_tile_config(128, 1, 0)
        bh.reset_kernel_times()
        t1 = time.perf_counter()
        bh.ls_tiles_device(*args_)
        bh.synchronize()
        ms64 = (time.perf_counter() - t1) * 1000
        f64 = d.to_host()
    finally:
        bh.set_tile_precision(bh.PREC_FP32)
        bh.set_tile_config(block, cpl, 0)
        bh.set_tile_lanes(lanes)
    if keep is not None:
        keep.update(tiles=tiles, w=w, n_pts=n_pts, f64=f64)
    flop = TILE_STEP_FLOP * nsteps * steps
    ach = flop / (kms / 1000) / 1e12
    cells = nt * w * w
    return {"ms_per_map": ms, "value": cells / (ms / 1000), "integration_steps_per_map": nsteps,
            "autotuner": {"block": block, "cells_per_lane": cpl, "lanes_per_cell": lanes,
                          "candidates_block_lanes_cpl_ms": tune},
            "fp64_tiles_ms": ms64, "fp32_vs_fp64": tiles_fp32_vs_fp64(bh, np, f32, f64),
            "roofline": {"bound": "valu_fp32", "kernel": "k_ls_tiles", "achieved": ach,
                         "peak": FP32_VALU_PEAK / 1e12, "unit": "TFLOP/s",
                         "frac": ach / (FP32_VALU_PEAK / 1e12), "traffic": None,
                         "avg_launch_ms": kms / max(nl, 1), "launches": nl,
                         "note": "48 FLOP per step (state + 3 tangent vectors, FMA = 2); built "
                                 "without packed FP32 instructions"}}


def run_tiles(bh, np, args, out):
    """Lyapunov spectrum of uneven refinement tiles (config 4), with the
    FP32 spectrum's error against the FP64 tile kernel."""
    keep = {}
    r = run_config4(bh, np, args.steps, args.warmup, keep)
    out.update(metric="grid points/s, uneven-tile Lyapunov spectrum (FP32 state, FP64 GS)",
               value=r["value"], unit="grid points/s", ms_per_step=r["ms_per_map"], dtype="f32",
               config={"workload": "4096 tiles of 16x16 cells on the Rossler 4096x4096 (a, c) "
                                   f"lattice, lengths U[200, 2000] t.u., NT {LLE['NT']}",
                       "integration_steps_per_map": r["integration_steps_per_map"]})
    out["autotuner"] = r["autotuner"]
    out["fp64_tiles_ms"] = r["fp64_tiles_ms"]
    out["fp32_vs_fp64"] = r["fp32_vs_fp64"]
    out["roofline"] = r["roofline"]
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = tiles_cpu_baseline(np, keep["tiles"], keep["w"], keep["n_pts"],
                                                 keep["f64"])


LLE32_STEP_FLOP = 11         # FP32 fused-multiply-add form of the Rossler step (FMA = 2)
LLE32_DEV_FLOP = 13          # FP32 deviation map of the perturbed trajectory (FMA = 2)


def run_lle(bh, np, args, out):
    """2D LLE map of the Rossler box on an n x n lattice: FP32 bulk
    integration (FP64 norm, log accumulation and renormalisation) timed
    against the plain FP64 map of the same lattice, with the FP32 map's
    error and sign / regime agreement against the FP64 run.  Perturbation
    1e-8 along x."""
    n = args.n or 1024
    cells = n * n
    lle = dict(LLE, NT=args.nt)
    d64 = bh.DeviceArray((cells,), np.float64)
    d32 = bh.DeviceArray((cells,), np.float64)
    lle_args = (2, lle["tMax"], lle["NT"], n, H, lle["eps"], IC, RANGES, [0, 2], 0.0,
                lle["transient"], VALS)

    def timed(dst, prec, steps, kernel):
        bh.set_precision(prec)
        for _ in range(args.warmup):
            bh.lle_map_device(dst.ptr, *lle_args)
        bh.synchronize()
        bh.set_timing(True)
        bh.reset_kernel_times()
        t0 = time.perf_counter()
        for _ in range(steps):
            bh.lle_map_device(dst.ptr, *lle_args)
        bh.synchronize()
        ms = (time.perf_counter() - t0) * 1000 / steps
        kms, nl = bh.kernel_time(kernel)
        return ms, kms, nl

    ms64, k64, n64 = timed(d64, bh.PREC_FP64, 1, "lle")
    ms32, k32, n32 = timed(d32, bh.PREC_FP32, args.steps, "lle32")
    bh.set_precision(bh.PREC_FP64)
    a, b = d64.to_host(), d32.to_host()
    fin = np.isfinite(a) & np.isfinite(b)
    err = np.abs(a[fin] - b[fin])
    with np.errstate(invalid="ignore"):
        sign = float((np.sign(a[fin]) == np.sign(b[fin])).mean())
        chaos = float(((a > CHAOS) == (b > CHAOS)).mean())
    skip = int(lle["transient"] / H)
    nNT = int(lle["NT"] / H)
    nb = int(lle["tMax"] / lle["NT"])
    flop32 = (LLE32_STEP_FLOP * (skip + nb * nNT) + LLE32_DEV_FLOP * nb * nNT) * cells
    ach = flop32 * n32 / (k32 / 1000) / 1e12
    out.update(metric="grid points/s, 2D LLE map, FP32 bulk vs FP64", value=cells / (ms32 / 1000),
               unit="grid points/s", ms_per_step=ms32, dtype="f32",
               config={"workload": f"2D LLE map, Rossler {n}x{n} (a, c) lattice, FP32 bulk "
                                   f"integration with FP64 norm / log / renormalisation, "
                                   f"perturbation 1e-8 along x",
                       "grid": n, "ranges": RANGES, "h": H, "lle": lle})
    out["fp64_reference"] = {"ms": ms64, "value": cells / (ms64 / 1000),
                             "kernel_ms": k64 / max(n64, 1), "speedup_fp32_vs_fp64": ms64 / ms32}
    out["error_vs_fp64"] = {"cells_compared": int(fin.sum()),
                            "nan_cells_agree": bool(np.array_equal(np.isnan(a), np.isnan(b))),
                            "max_abs": float(err.max()), "mean_abs": float(err.mean()),
                            "p99_abs": float(np.percentile(err, 99)),
                            "sign_agreement": sign, "chaos_agreement": chaos,
                            "chaos_threshold": CHAOS}
    out["roofline"] = {"bound": "valu_fp32", "kernel": "k_lle32", "achieved": ach,
                       "peak": FP32_VALU_PEAK / 1e12, "unit": "TFLOP/s",
                       "frac": ach / (FP32_VALU_PEAK / 1e12), "traffic": None,
                       "avg_launch_ms": k32 / max(n32, 1), "launches": n32,
                       "note": f"{LLE32_STEP_FLOP} FLOP per state step, {LLE32_DEV_FLOP} per "
                               f"deviation step (FMA = 2); built without packed FP32 "
                               f"instructions"}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = lle_cpu_baseline(np, n, lle, a)


def run_config3(bh, np):
    """Config 3 in brief: the 1024^2 LLE map in FP64 and with FP32 bulk
    integration (transient 500, measure 1000, NT 0.5, perturbation 1e-8
    along x), one map each, and the FP32 map's agreement with the FP64 one."""
    n = 1024
    d64 = bh.DeviceArray((n * n,), np.float64)
    d32 = bh.DeviceArray((n * n,), np.float64)
    la = (2, LLE["tMax"], LLE["NT"], n, H, LLE["eps"], IC, RANGES, [0, 2], 0.0,
          LLE["transient"], VALS)
    res = {}
    try:
        for name, d, prec in (("fp64", d64, bh.PREC_FP64), ("fp32_bulk", d32, bh.PREC_FP32)):
            bh.set_precision(prec)
            bh.lle_map_device(d.ptr, *la)     # warm-up
            bh.synchronize()
            t0 = time.perf_counter()
            bh.lle_map_device(d.ptr, *la)
            bh.synchronize()
            res[name + "_ms"] = (time.perf_counter() - t0) * 1000
    finally:
        bh.set_precision(bh.PREC_FP64)
    a, b = d64.to_host(), d32.to_host()
    ok = np.isfinite(a) & np.isfinite(b)
    err = np.abs(a[ok] - b[ok])
    with np.errstate(invalid="ignore"):
        res.update(regime_agreement=float(((a > CHAOS) == (b > CHAOS)).mean()),
                   sign_agreement=float((np.sign(a[ok]) == np.sign(b[ok])).mean()),
                   nan_in_one_only=float((np.isnan(a) != np.isnan(b)).mean()),
                   lle_abs_err_mean=float(err.mean()),
                   lle_abs_err_p99=float(np.quantile(err, 0.99)),
                   lle_abs_err_max=float(err.max()))
    res["speedup_fp32_vs_fp64"] = res["fp64_ms"] / res["fp32_bulk_ms"]
    return res


BIF_FLOP_PER_CELL = STEP_FLOP * int((PEAK["transient"] + PEAK["tMax"]) / H)   # 1.8 MFLOP


def run_dbscan(bh, np, args, out):
    """2D bifurcation map: DBSCAN cluster count of the (amplitude, interval)
    peak pairs of x per cell, plain FP64, Rossler n x n over the (a, c) box,
    transient 1000, record 500, eps 0.01, counts capped at 32 (the
    reference's maxAmountOfPeaks is a threshold, so it is set to the record
    length and the cap is the separate clamp).  The CPU baseline leg times
    the oracle on three rows and compares them with the map bit for bit."""
    n = args.n or 512
    cells = n * n
    d = bh.DeviceArray((cells,), np.int32)
    bh.set_cluster_clamp(32)
    dargs = (d.ptr, PEAK["tMax"], n, H, IC, RANGES, [0, 2], 0, 0.0, n_rec(), PEAK["transient"],
             VALS, 1, PEAK["eps"])
    for _ in range(args.warmup):
        bh.dbscan_map_device(*dargs)
    bh.synchronize()
    bh.set_timing(True)
    bh.reset_kernel_times()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bh.dbscan_map_device(*dargs)
    bh.synchronize()
    ms = (time.perf_counter() - t0) * 1000 / args.steps
    fam = "bif_peaks2" if bh.kernel_time("bif_peaks2")[1] else "bif_peaks"
    kms, nl = bh.kernel_time(fam)
    dbms, _ = bh.kernel_time("dbscan")
    got = d.to_host().reshape(n, n)
    bh.set_cluster_clamp(0)
    ach = BIF_FLOP_PER_CELL * cells * args.steps / (kms / 1000) / 1e12
    out.update(metric="grid points/s, 2D DBSCAN peak-count map (plain FP64)",
               value=cells / (ms / 1000), unit="grid points/s", ms_per_step=ms, dtype="f64",
               config={"workload": f"2D bifurcation map, Rossler {n}x{n} (a, c) lattice, DBSCAN "
                                   f"peak count of x capped at 32, plain FP64",
                       "grid": n, "ranges": RANGES, "h": H, "peak": PEAK, "cap": 32})
    out["kernels_ms_per_step"] = {fam: kms / args.steps, "dbscan": dbms / args.steps}
    if args.fp32:
        out["fp32"] = run_dbscan_fp32(bh, np, args, d, dargs, got, ms)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = dbscan_cpu_baseline(np, n, got)
    out["roofline"] = {"bound": "valu_fp64", "kernel": "k_" + fam, "achieved": ach,
                       "peak": FP64_VALU_PEAK / 1e12, "unit": "TFLOP/s",
                       "frac": ach / (FP64_VALU_PEAK / 1e12), "traffic": None,
                       "avg_launch_ms": kms / max(nl, 1), "launches": nl,
                       "note": "12 FLOP per Rossler step, 150 000 steps per cell; no-contraction "
                               "FP64 reaches at most 50% of the FMA-counted peak"}


KDE = {"sampling": 10, "I1": -15.0, "I2": 15.0, "smooth": 0.05}


def run_kde(bh, np, args, out):
    """2D bifurcation map: KDE peak count of the x peaks per cell, plain
    FP64 trajectories, Rossler n x n over the (a, c) box, transient 1000,
    record 500, kdeSampling 10 over [-15, 15], smoothing 0.05.  The CPU
    baseline leg times the oracle on three rows and states their count
    agreement with the map (the oracle's expf is glibc's, so KDE counts are
    compared as agreement)."""
    n = args.n or 512
    cells = n * n
    d = bh.DeviceArray((cells,), np.int32)
    kargs = (d.ptr, PEAK["tMax"], n, H, IC, RANGES, [0, 2], 0, 0.0, n_rec(), PEAK["transient"],
             VALS, 1, KDE["sampling"], KDE["I1"], KDE["I2"], KDE["smooth"])
    for _ in range(args.warmup):
        bh.kde_map_device(*kargs)
    bh.synchronize()
    bh.set_timing(True)
    bh.reset_kernel_times()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bh.kde_map_device(*kargs)
    bh.synchronize()
    ms = (time.perf_counter() - t0) * 1000 / args.steps
    fam = "bif_peaks2" if bh.kernel_time("bif_peaks2")[1] else "bif_peaks"
    kms, nl = bh.kernel_time(fam)
    kdems, _ = bh.kernel_time("kde")
    got = d.to_host().reshape(n, n)
    ach = BIF_FLOP_PER_CELL * cells * args.steps / (kms / 1000) / 1e12
    out.update(metric="grid points/s, 2D KDE peak-count map (plain FP64)",
               value=cells / (ms / 1000), unit="grid points/s", ms_per_step=ms, dtype="f64",
               config={"workload": f"2D bifurcation map, Rossler {n}x{n} (a, c) lattice, KDE "
                                   f"peak count of x, plain FP64",
                       "grid": n, "ranges": RANGES, "h": H, "peak": PEAK, "kde": KDE})
    out["kernels_ms_per_step"] = {fam: kms / args.steps, "kde": kdems / args.steps}
    if not args.no_cpu_baseline:
        import orclib as O
        O.lib()
        rows = [0, n // 3, n - 1]
        t0 = time.perf_counter()
        wants = [O.kde_cells(PEAK["tMax"], n, H, IC, RANGES, [0, 2], 0, 0.0, n_rec(),
                             PEAK["transient"], VALS, 1, KDE["sampling"], KDE["I1"], KDE["I2"],
                             KDE["smooth"], r * n, n) for r in rows]
        dt = time.perf_counter() - t0
        agree = float(np.mean([np.mean(got[r] == w) for r, w in zip(rows, wants)]))
        out["cpu_baseline"] = {"value": len(rows) * n / dt, "unit": "grid points/s",
                               "cores": O.lib().orc_threads(), "cpu_model": cpu_model(),
                               "kind": "port", "sample": f"rows {rows} of the {n}x{n} map, "
                                                         f"{dt:.2f} s",
                               "rows_count_agreement_vs_gpu": agree}
    out["roofline"] = {"bound": "valu_fp64", "kernel": "k_" + fam, "achieved": ach,
                       "peak": FP64_VALU_PEAK / 1e12, "unit": "TFLOP/s",
                       "frac": ach / (FP64_VALU_PEAK / 1e12), "traffic": None,
                       "avg_launch_ms": kms / max(nl, 1), "launches": nl,
                       "note": "12 FLOP per Rossler step, 150 000 steps per cell; the KDE pass "
                               "skips the terms that are exactly +0 (peaks beyond 15 smoothing "
                               "widths of every sample point of a 64-point batch)"}


def dbscan_cpu_baseline(np, n, got):
    """CPU baseline of the DBSCAN map: the oracle (OpenMP over cells) on
    three whole rows, timed, and its counts compared with the GPU map's."""
    import orclib as O
    O.lib()
    rows = [0, n // 3, n - 1]
    exact = True
    t0 = time.perf_counter()
    wants = [O.dbscan_cells(PEAK["tMax"], n, H, IC, RANGES, [0, 2], 0, 0.0, n_rec(),
                            PEAK["transient"], VALS, 1, PEAK["eps"], r * n, n) for r in rows]
    dt = time.perf_counter() - t0
    for r, want in zip(rows, wants):
        exact &= bool(np.array_equal(got[r], np.minimum(want, 32)))
    return {"value": len(rows) * n / dt, "unit": "grid points/s", "cores": O.lib().orc_threads(),
            "cpu_model": cpu_model(), "kind": "port",
            "sample": f"rows {rows} of the {n}x{n} map, {dt:.2f} s",
            "rows_bit_exact_vs_gpu": exact}


def lle_cpu_baseline(np, n, lle, got64, rows=32, repeats=3):
    """CPU baseline of the 2D LLE map: the oracle (FP64, OpenMP over cells,
    two trajectories per cell as the reference steps them) on `rows` evenly
    spaced whole rows, median of `repeats` runs, and its values compared
    with the GPU FP64 map's rows (NaN cells as a class, else <= 1e-12)."""
    import orclib as O
    O.lib()
    sel = list(range(0, n, max(1, n // rows)))[:rows]
    times, want = [], None
    for _ in range(repeats):
        t0 = time.perf_counter()
        want = [O.lle_cells(2, lle["tMax"], lle["NT"], n, H, lle["eps"], IC, RANGES, [0, 2],
                            0.0, lle["transient"], VALS, O.PERT_AXIS_X, 0, r * n, n)
                for r in sel]
        times.append(time.perf_counter() - t0)
    got = got64.reshape(n, n)
    nan_ok, err = True, 0.0
    for r, w in zip(sel, want):
        gn, wn = np.isnan(got[r]), np.isnan(w)
        nan_ok &= bool(np.array_equal(gn, wn))
        if (~wn).any():
            err = max(err, float(np.abs(got[r][~wn] - w[~wn]).max()))
    cells = len(sel) * n
    dt = sorted(times)[len(times) // 2]
    return {"value": cells / dt, "unit": "grid points/s", "cores": O.lib().orc_threads(),
            "cpu_model": cpu_model(), "kind": "port",
            "build": "gcc -O3 -ffp-contract=off -fopenmp", "runs_s": times,
            "sample": f"{cells} cells ({len(sel)} evenly spaced rows of the {n}x{n} lattice), "
                      f"FP64 LLE per cell, median of {repeats} runs: {dt:.1f} s",
            "fp64_rows_vs_gpu": {"nan_cells_agree": nan_ok, "max_abs_err": err}}


def tiles_cpu_baseline(np, tiles, w, n_pts, f64, ntiles=128, repeats=3):
    """CPU baseline of the uneven tiles: the FP64 variational oracle (the
    reference's Gram-Schmidt every 0.5 time units, OpenMP over cells) on
    `ntiles` tiles spread over the length range, median of `repeats` runs,
    and its spectra compared with the GPU FP64 tile kernel's."""
    import orclib as O
    O.lib()
    oi, oj, nb = tiles
    nt = len(nb)
    order = np.argsort(nb)
    sel = order[np.linspace(0, nt - 1, ntiles).astype(int)]
    a, c, blocks = [], [], []
    for t in sel:
        for cell in range(w * w):
            i, j = oi[t] + cell % w, oj[t] + cell // w
            a.append(O.value_by_idx(j * n_pts + i, n_pts, RANGES[0], RANGES[1], 0))
            c.append(O.value_by_idx(j * n_pts + i, n_pts, RANGES[2], RANGES[3], 1))
            blocks.append(nb[t])
    nNT = int(LLE["NT"] / H)
    times, want = [], None
    for _ in range(repeats):
        t0 = time.perf_counter()
        want = O.ls_var_cells(a, [VALS[1]] * len(a), c, IC, H, 0, nNT, blocks)
        times.append(time.perf_counter() - t0)
    got = f64.reshape(nt, w * w, 3)[sel].reshape(-1, 3)
    fo = np.isfinite(want).all(axis=1)
    steps = int(np.asarray(blocks, dtype=np.int64).sum()) * nNT
    dt = sorted(times)[len(times) // 2]
    return {"value": len(a) / dt, "unit": "grid points/s", "cores": O.lib().orc_threads(),
            "cpu_model": cpu_model(), "kind": "port",
            "build": "gcc -O3 -ffp-contract=off -fopenmp", "runs_s": times,
            "sample": f"{len(a)} cells ({ntiles} tiles spread over the length range, "
                      f"{steps} integration steps), FP64 variational spectrum, median of "
                      f"{repeats} runs: {dt:.1f} s",
            "fp64_tiles_vs_gpu": {
                "finite_cells_agree": bool(np.array_equal(np.isfinite(got).all(axis=1), fo)),
                "max_abs_err": float(np.abs(got[fo] - want[fo]).max()) if fo.any() else 0.0}}


def run_dbscan_fp32(bh, np, args, d, dargs, got64, ms64):
    """The same map with FP32 trajectories (bifhip_set_precision FP32):
    time and peak-count agreement with the FP64 map (bit-exactness against
    the oracle's FP32 restatement is tests/test_gpu_fp32_peaks.py's)."""
    n = int(round(got64.size ** 0.5))
    cells = n * n
    bh.set_precision(bh.PREC_FP32)
    bh.set_cluster_clamp(32)
    try:
        for _ in range(args.warmup):
            bh.dbscan_map_device(*dargs)
        bh.synchronize()
        bh.reset_kernel_times()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            bh.dbscan_map_device(*dargs)
        bh.synchronize()
        ms = (time.perf_counter() - t0) * 1000 / args.steps
        kms, nl = bh.kernel_time("bif_peaks")
        dbms, _ = bh.kernel_time("dbscan")
        got = d.to_host().reshape(n, n)
    finally:
        bh.set_precision(bh.PREC_FP64)
        bh.set_cluster_clamp(0)
    c64, c32 = np.clip(got64, 1, 16), np.clip(got, 1, 16)
    # 11 FLOP per FP32 step (4 fused multiply-adds, 3 adds / multiplies)
    ach = 11 * 150000 * cells * args.steps / (kms / 1000) / 1e12
    return {"ms_per_step": ms, "value": cells / (ms / 1000), "speedup_vs_fp64": ms64 / ms,
            "count_agreement_vs_fp64": float((got == got64).mean()),
            "class_agreement_vs_fp64": float((c32 == c64).mean()),
            "kernels_ms_per_step": {"bif_peaks32": kms / args.steps, "dbscan": dbms / args.steps},
            "roofline": {"bound": "valu_fp32", "kernel": "k_bif_peaks32", "achieved": ach,
                         "peak": FP32_VALU_PEAK / 1e12, "unit": "TFLOP/s",
                         "frac": ach / (FP32_VALU_PEAK / 1e12), "launches": nl}}


def run_drivers(bh, np, args, out):
    """The reference's file-writing drivers (Library/hostLibrary.cuh), each
    called as a drop-in caller would at its configuration's size: host
    inputs in, the CSV file written.  Per driver: wall time per call, the
    kernel time inside it and the file size.  The line's value is the LLE2D
    driver's grid points/s (1024^2, reference perturbation directions)."""
    import os
    import tempfile
    ic3 = (3, IC)
    lle_common = (LLE["tMax"], LLE["NT"])
    drivers = {
        "bifurcation1D": (bh.bifurcation1D, 1000, ("bif_peaks",),
                          (PEAK["tMax"], 1000, H, 3, IC, [2.0, 7.0], [2], 0, 0.0,
                           PEAK["transient"], [0.2, 0.2, 0.0], 3, 1)),
        "bifurcation2DDBSCAN": (bh.bifurcation2DDBSCAN, 512 * 512,
                                ("bif_peaks", "bif_peaks2", "dbscan"),
                                (PEAK["tMax"], 512, H, *ic3, RANGES, [0, 2], 0, 0.0, n_rec(),
                                 PEAK["transient"], VALS, 3, 1, PEAK["eps"])),
        "bifurcation2DKDE": (bh.bifurcation2DKDE, 512 * 512, ("bif_peaks", "bif_peaks2", "kde"),
                             (PEAK["tMax"], 512, H, *ic3, RANGES, [0, 2], 0, 0.0, n_rec(),
                              PEAK["transient"], VALS, 3, 1, KDE["sampling"], KDE["I1"],
                              KDE["I2"], KDE["smooth"])),
        "LLE1D": (bh.LLE1D, 1000, ("lle",),
                  (*lle_common, 1000, H, LLE["eps"], IC, 3, [2.0, 7.0], [2], 0, 0.0,
                   LLE["transient"], [0.2, 0.2, 0.0], 3)),
        "LLE2D": (bh.LLE2D, 1024 * 1024, ("lle",),
                  (*lle_common, 1024, H, LLE["eps"], IC, 3, RANGES, [0, 2], 0, 0.0,
                   LLE["transient"], VALS, 3)),
        "LS1D": (bh.LS1D, 1000, ("ls",),
                 (*lle_common, 1000, H, 1e-6, IC, 3, [2.0, 7.0], [2], 0, 0.0,
                  LLE["transient"], [0.2, 0.2, 0.0], 3)),
        "LS2D": (bh.LS2D, 512 * 512, ("ls",),
                 (*lle_common, 512, H, 1e-6, IC, 3, RANGES, [0, 2], 0, 0.0,
                  LLE["transient"], VALS, 3)),
    }
    res = {}
    bh.set_perturbation(bh.PERT_REFERENCE)       # the drivers' own default
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "mat.csv")
        bh.set_output_path(path)
        bh.set_timing(True)
        for name, (fn, cells, fams, dargs) in drivers.items():
            for f in os.listdir(tmp):
                os.remove(os.path.join(tmp, f))
            fn(*dargs)                                     # warm-up
            bh.reset_kernel_times()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                fn(*dargs)
            ms = (time.perf_counter() - t0) * 1000 / args.steps
            kms = sum(bh.kernel_time(f)[0] for f in fams) / args.steps
            res[name] = {"cells": cells, "ms_per_call": ms, "kernel_ms": kms,
                         "host_ms": ms - kms,
                         "file_mb": sum(os.path.getsize(os.path.join(tmp, f))
                                        for f in os.listdir(tmp)) / 1e6,
                         "grid_points_per_s": cells / (ms / 1000)}
            print(f"driver {name}: {ms:.1f} ms per call, kernels {kms:.1f} ms", file=sys.stderr,
                  flush=True)
    lle = res["LLE2D"]
    out.update(metric="grid points/s, LLE2D through the reference's file-writing driver",
               value=lle["grid_points_per_s"], unit="grid points/s",
               ms_per_step=lle["ms_per_call"], dtype="f64",
               config={"workload": "the reference's drivers at their configuration sizes, host "
                                   "inputs, CSV files written; LLE2D on the Rossler 1024x1024 "
                                   "(a, c) lattice, transient 500, measure 1000, NT 0.5, "
                                   "reference perturbation directions"})
    out["drivers"] = res


def run_bif1d(bh, np, args, out):
    """1D bifurcation diagram: local maxima of x for c over [2, 7] in 1000
    points (a = b = 0.2), transient 1000, record 500, plain FP64 through the
    reference's driver with host buffers (the time includes the copy of the
    peak rows back to the host).  The CPU baseline leg times the oracle on
    every cell and compares its peaks with the GPU's."""
    n = 1000
    vals = [0.2, 0.2, 0.0]
    bargs = (PEAK["tMax"], n, H, IC, [2.0, 7.0], [2], 0, 0.0, PEAK["transient"], vals, 1)
    for _ in range(args.warmup):
        bh.bifurcation1D_to_buffer(*bargs)
    bh.set_timing(True)
    bh.reset_kernel_times()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cnt, pk = bh.bifurcation1D_to_buffer(*bargs)
    ms = (time.perf_counter() - t0) * 1000 / args.steps
    kms, _ = bh.kernel_time("bif_peaks")
    out.update(metric="parameter points/s, 1D bifurcation diagram (plain FP64)",
               value=n / (ms / 1000), unit="parameter points/s", ms_per_step=ms, dtype="f64",
               config={"workload": "1D bifurcation diagram, Rossler c in [2, 7] x 1000 points, "
                                   "a = b = 0.2, transient 1000, record 500, host buffers",
                       "h": H, "peak": PEAK})
    out["kernels_ms_per_step"] = {"bif_peaks": kms / args.steps}
    if not args.no_cpu_baseline:
        # CPU baseline leg: the oracle on every cell, timed; its peaks are
        # compared with the GPU diagram's
        import orclib as O
        O.lib()
        t1 = time.perf_counter()
        wc, wp, _ = O.peaks_cells(1, PEAK["tMax"], n, H, IC, [2.0, 7.0], [2], 0, 0.0,
                                  PEAK["transient"], vals, 1, 0, n, pk.shape[1])
        cpu_s = time.perf_counter() - t1
        exact = bool(np.array_equal(cnt, np.maximum(wc, 0))) and all(
            np.array_equal(pk[k, :cnt[k]], wp[k, :cnt[k]]) for k in range(n))
        out["cpu_baseline"] = {"value": n / cpu_s, "unit": "parameter points/s",
                               "cores": O.lib().orc_threads(), "cpu_model": cpu_model(),
                               "kind": "port", "sample": f"all {n} cells, {cpu_s:.2f} s",
                               "all_cells_bit_exact_vs_gpu": exact}


def run_copy(bh, np, args, out):
    """Device-to-device copy of a known size: the reference point that
    scales the FETCH_SIZE / WRITE_SIZE counters into bytes."""
    mib = 1024
    a = bh.DeviceArray((mib << 17,), np.float64)
    b = bh.DeviceArray((mib << 17,), np.float64)
    for _ in range(args.warmup + args.steps):
        bh._check(bh.load().bifhip_memcpy_d2d(b.ptr, a.ptr, mib << 20))
    bh.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bh._check(bh.load().bifhip_memcpy_d2d(b.ptr, a.ptr, mib << 20))
    bh.synchronize()
    ms = (time.perf_counter() - t0) * 1000 / args.steps
    out.update(metric="device copy GB/s (counter calibration)", value=2 * (mib << 20) / ms / 1e6,
               unit="GB/s", ms_per_step=ms, config={"workload": f"copy {mib} MiB"})


MFMA_BF16_PEAK = 2.5e15      # MI355X dense BF16 matrix peak
HBM_PEAK = 8.0e12            # MI355X HBM3E
MLP_TRAIN_FLOP_PER_ROW = 3 * 2 * 45312   # forward, input- and weight-gradient products
MLP_INFER_FLOP_PER_CELL = 86272
# DBSCAN reads each clustered cell's peak row once: about 85 (amplitude,
# interval) pairs per cell on this lattice (Rossler, 500-unit window; 85.3
# from the fused kernel's row writes), 16 B each, plus the count in and the
# result out
DBSCAN_BYTES_PER_CELL = 85 * 16 + 8
# HBM bytes per algorithmic byte, from the FETCH_SIZE / WRITE_SIZE counter
# runs scaled by the same counters on a 1 GiB copy (FETCH x 2.0, WRITE x
# 1.0): k_lle_peaks and k_bif_peaks2 writes, k_dbscan_wave reads
# (profiles/plain_fp64_lle_peak_map_1024.md, dbscan_peak_rows_1024.md)
TRAFFIC_RATIO = {"k_lle_peaks": 1.06, "k_bif_peaks2": 1.03, "k_dbscan_wave": 1.05}


def lle32_flop_per_cell():
    skip = int(LLE["transient"] / H)
    nNT = int(LLE["NT"] / H)
    nb = int(LLE["tMax"] / LLE["NT"])
    return LLE32_STEP_FLOP * (skip + nb * nNT) + LLE32_DEV_FLOP * nb * nNT


def agreement(np, peak, lle, ref):
    """Per-cell agreement of a map with the plain FP64 map under each reading
    of a cell's regime, and the LLE error where both are finite."""
    t_peak, t_lle, t_cls, t_ch = ref
    cls, ch = regimes(np, peak, lle)
    fin = np.isfinite(lle) & np.isfinite(t_lle)
    err = np.abs(lle[fin] - t_lle[fin])
    return {"regime": float(((cls == t_cls) & (ch == t_ch)).mean()),
            "chaos": float((ch == t_ch).mean()),
            "class": float((cls == t_cls).mean()),
            "count_exact": float((peak == t_peak).mean()),
            "lle_abs_err_mean": float(err.mean()) if err.size else 0.0,
            "lle_abs_err_p99": float(np.percentile(err, 99)) if err.size else 0.0,
            "lle_abs_err_max": float(err.max()) if err.size else 0.0}


def summarise(np, per_step, key):
    v = np.array([a[key] for a in per_step])
    return {"mean": float(v.mean()), "min": float(v.min()), "max": float(v.max())}


def noise_ceiling(np, ref, n, plain_ms, fixed_ms, cost):
    """What the plain map's own cell-to-cell noise allows.  A predictor that
    knows the exact regimes of a cell's 8 neighbours and takes their
    majority is the best local predictor; the cells it gets wrong beyond
    the 1% budget must be re-solved by any method that reaches 99%, at no
    less than the cost of the cheapest re-solve that settles the reading
    (ms per cell: the peak count alone for a class error, the FP32 bulk LLE
    for a chaos-bit error), the 1% budget spent on the dearest cells.  The
    bound is plain_ms / (fixed cost of the network + those re-solves)."""
    _, _, t_cls, t_ch = ref
    c = t_cls.reshape(n, n)
    h = t_ch.reshape(n, n).astype(np.int8)
    pad_c = np.pad(c, 1, mode="edge")
    pad_h = np.pad(h, 1, mode="edge")
    nb_c = [pad_c[1 + di:1 + di + n, 1 + dj:1 + dj + n]
            for di in (-1, 0, 1) for dj in (-1, 0, 1) if di or dj]
    nb_h = sum(pad_h[1 + di:1 + di + n, 1 + dj:1 + dj + n]
               for di in (-1, 0, 1) for dj in (-1, 0, 1) if di or dj)
    stack = np.stack(nb_c)
    votes = np.stack([(stack == k).sum(axis=0) for k in range(1, 17)])
    maj_c = votes.argmax(axis=0) + 1
    maj_h = nb_h >= 4
    wc = maj_c != c
    wh = maj_h != h.astype(bool)
    err_c = float(wc.mean())
    err_h = float(wh.mean())
    err_r = float((wc | wh).mean())

    def bound(groups):
        # groups: (fraction of cells, ms per cell to settle them); the 1%
        # budget is spent on the dearest cells first
        left, ms = 0.01, fixed_ms
        for f, c_ in sorted(groups, key=lambda g: -g[1]):
            skip = min(left, f)
            left -= skip
            ms += (f - skip) * n * n * c_
        return plain_ms / ms

    only_h = float((wh & ~wc).mean())

    return {"reference_boundary_fraction": boundary_fraction(np, t_cls, t_ch, n),
            "class_boundary_fraction": boundary_fraction(np, t_cls, np.zeros_like(t_ch), n),
            "chaos_boundary_fraction": boundary_fraction(np, np.ones_like(t_cls), t_ch, n),
            "neighbour_majority_error": {"regime": err_r, "class": err_c, "chaos": err_h},
            "speedup_bound_at_99": {"regime": bound([(err_c, cost["class"]),
                                                     (only_h, cost["chaos"])]),
                                    "class": bound([(err_c, cost["class"])]),
                                    "chaos": bound([(err_h, cost["chaos"])])},
            "resolve_ms_per_cell": cost,
            "bound_model": "plain_ms / (network ms + (neighbour-majority error - 0.01) x cells "
                           "x cheapest re-solve ms per cell that settles the error)"}


def kernel_rooflines(bh, args, stats, n, step_ms):
    """Every kernel family of the timed surrogate steps: time per step, share
    of the step, algorithmic work (the cells each family's launches covered,
    counted by the library, times the work per cell) / HIP-event time
    against its peak.  DBSCAN runs on a side stream beside the integrators,
    so its share overlaps theirs."""
    K = len(stats)
    steps_per_epoch = args.coarse ** 2 // min(args.batch, args.coarse ** 2)
    train_rows = steps_per_epoch * min(args.batch, args.coarse ** 2) * 20 * K
    rows = [
        ("k_lle_peaks", "lle_peaks", "valu_fp64", lle_flop_per_cell(), FP64_VALU_PEAK),
        ("k_bif_peaks2", "bif_peaks2", "valu_fp64", BIF_FLOP_PER_CELL, FP64_VALU_PEAK),
        ("k_bif_peaks", "bif_peaks", "valu_fp64", BIF_FLOP_PER_CELL, FP64_VALU_PEAK),
        ("k_lle32", "lle32", "valu_fp32", lle32_flop_per_cell(), FP32_VALU_PEAK),
        ("k_lle64", "lle", "valu_fp64", lle_flop_per_cell(), FP64_VALU_PEAK),
        ("k_dbscan_wave", "dbscan", "hbm", DBSCAN_BYTES_PER_CELL, HBM_PEAK),
        ("k_train_fused", "mlp_train", "mfma_bf16", None, MFMA_BF16_PEAK),
        ("k_infer_fused", "mlp_infer", "mfma_bf16", None, MFMA_BF16_PEAK),
    ]
    out = {}
    for name, timer, bound, per_cell, peak in rows:
        ms, launches = bh.kernel_time(timer)
        if launches == 0:
            continue
        cells = bh.kernel_cells(timer)
        if name == "k_train_fused":
            work = MLP_TRAIN_FLOP_PER_ROW * train_rows
        elif name == "k_infer_fused":
            work = MLP_INFER_FLOP_PER_CELL * n * n * K
        else:
            work = per_cell * cells
        ach = work / (ms / 1000) if ms > 0 else 0.0
        hbm = bound == "hbm"
        row = {"bound": bound, "ms_per_step": ms / K, "share_of_step": ms / K / step_ms,
               "launches": launches, "avg_launch_ms": ms / launches,
               "achieved": ach / (1e9 if hbm else 1e12),
               "peak": peak / (1e9 if hbm else 1e12),
               "unit": "GB/s" if hbm else "TFLOP/s", "frac": ach / peak,
               "work_per_launch": work / launches, "cells": cells,
               "traffic_ratio": TRAFFIC_RATIO.get(name)}
        # HBM bytes written per launch by the peak kernels: counter ratio x
        # algorithmic peak-row bytes
        if name == "k_lle_peaks":
            row["traffic"] = LLE_PEAKS_HBM_BYTES_PER_CELL * cells / launches
        elif name == "k_bif_peaks2":
            row["traffic"] = TRAFFIC_RATIO[name] * (DBSCAN_BYTES_PER_CELL - 4) * cells / launches
        out[name] = row
    if "k_dbscan_wave" in out:
        out["k_dbscan_wave"]["note"] = ("side stream, overlapped with the integrators; the "
                                        "family includes the rare workgroup-kernel cells; bound "
                                        "by the label machine's dependent compare and mask "
                                        "steps, not by HBM")
    if "k_train_fused" in out:
        blocks = -(-min(args.batch, args.coarse ** 2) // 64)
        out["k_train_fused"]["note"] = (
            f"BF16 MFMA, latency-bound: each step is {blocks} workgroups of 64 rows, one "
            f"dependent layer after another; training is "
            f"{100 * out['k_train_fused']['share_of_step']:.1f}% of the step")
    if "k_infer_fused" in out:
        out["k_infer_fused"]["note"] = (
            f"BF16 MFMA over every fine cell; "
            f"{100 * out['k_infer_fused']['share_of_step']:.1f}% of the step")
    if "k_bif_peaks" in out:
        out["k_bif_peaks"]["note"] = "one chain per lane: recounts and overflow recomputes"
    return out


def run_surrogate(bh, np, args, out, n, d_peak, d_lle):
    """The headline: surrogate-refined 4096^2 map against the plain FP64 map
    of the same lattice.  The plain map is warmed up at full size and timed
    as the mean of --plain-steps runs.  Every timed surrogate step writes
    its own output buffers, so agreement is scored on each of them (seeds
    seed_base .. seed_base + steps - 1); the line is marked invalid, and the
    process exits with status 3, when the mean regime agreement is below
    0.99.  Variants timed on the first --variant-steps seeds: every regime
    re-solve in FP64, and the chaos-bit-only and class-only readings of a
    cell's regime."""
    cells = n * n
    plain_map(bh, n, d_peak, d_lle)               # warm-up at full size
    bh.synchronize()
    plain = []
    for _ in range(max(1, args.plain_steps)):
        t0 = time.perf_counter()
        plain_map(bh, n, d_peak, d_lle)
        bh.synchronize()
        plain.append((time.perf_counter() - t0) * 1000)
    plain_ms = float(np.mean(plain))
    ref_peak, ref_lle = d_peak.to_host(), d_lle.to_host()
    t_cls, t_ch = regimes(np, ref_peak, ref_lle)
    ref = (ref_peak, ref_lle, t_cls, t_ch)

    K = args.steps
    bufs = [(bh.DeviceArray((cells,), np.int32), bh.DeviceArray((cells,), np.float64),
             bh.DeviceArray((cells,), np.uint8)) for _ in range(max(K, args.variant_steps, 1))]

    def run(k, seed0, warm, validation=None, batch=None):
        validation = args.validation if validation is None else validation
        batch = args.batch if batch is None else batch
        for w in range(warm):
            surrogate_map(bh, n, args.coarse, *bufs[0], seed0 + 100 + w, args.conf,
                          validation, batch)
        bh.synchronize()
        bh.reset_kernel_times()
        st = []
        t0 = time.perf_counter()
        for s in range(k):
            st.append(surrogate_map(bh, n, args.coarse, *bufs[s], seed0 + s, args.conf,
                                    validation, batch))
        bh.synchronize()
        return (time.perf_counter() - t0) * 1000 / max(k, 1), st

    def scored(k):
        return [agreement(np, bufs[s][0].to_host(), bufs[s][1].to_host(), ref) for s in range(k)]

    bh.set_timing(True)
    ms, stats = run(K, args.seed_base, args.warmup)
    rf = kernel_rooflines(bh, args, stats, n, ms)
    per = scored(K)
    agree = summarise(np, per, "regime")
    last = stats[-1]
    res = bufs[K - 1][2].to_host()
    s_lle = bufs[K - 1][1].to_host()
    s_ch = regimes(np, bufs[K - 1][0].to_host(), s_lle)[1]
    k32 = (res == 6) | (res == 7)
    e32 = np.abs(s_lle[k32] - ref_lle[k32])
    e32 = e32[np.isfinite(e32)]
    fp32_lle = {"cells": int(k32.sum()),
                "lle_abs_err_mean": float(e32.mean()) if e32.size else 0.0,
                "lle_abs_err_max": float(e32.max()) if e32.size else 0.0,
                "regime_agreement": float((s_ch[k32] == t_ch[k32]).mean()) if k32.any() else 1.0}
    fixed_ms = float(np.mean([st["ms_coarse"] + st["ms_train"] + st["ms_infer"] + st["ms_flag"]
                              for st in stats]))
    peak_ms = bh.kernel_time("bif_peaks2")[0] + bh.kernel_time("bif_peaks")[0]
    peak_cells = bh.kernel_cells("bif_peaks2") + bh.kernel_cells("bif_peaks")
    lle32_ms, _ = bh.kernel_time("lle32")
    lle32_cells = bh.kernel_cells("lle32")
    cost = {"class": peak_ms / max(peak_cells, 1)}
    cost["chaos"] = lle32_ms / lle32_cells if lle32_cells and lle32_ms > 0 else cost["class"]

    variants = {}
    kv = max(1, min(args.variant_steps, len(bufs)))
    def breakdown(k, vst):
        fams = ("lle_peaks", "bif_peaks2", "bif_peaks", "lle32", "lle", "dbscan", "mlp_train",
                "mlp_infer")
        kern = {f: {"ms_per_step": bh.kernel_time(f)[0] / k, "cells_per_step": bh.kernel_cells(f) / k}
                for f in fams if bh.kernel_time(f)[1]}
        phases = {p_: float(np.mean([st[p_] for st in vst]))
                  for p_ in ("ms_coarse", "ms_train", "ms_infer", "ms_flag", "ms_solve")}
        return {"kernels": kern, "phase_ms": phases}

    if args.variant_steps > 0:
        try:
            bh.set_lle32_margin(-1.0)
            vms, vst = run(kv, args.seed_base, 1)
            variants["all_fp64_resolves"] = {
                "ms_per_step": vms, "speedup_vs_plain_fp64": plain_ms / vms, "steps": kv,
                **breakdown(kv, vst),
                "agreement": {k: summarise(np, scored(kv), k) for k in ("regime", "chaos",
                                                                         "class")},
                "note": "every regime re-solve in FP64 (lle32 margin < 0)"}
        finally:
            bh.set_lle32_margin(args.lle32_margin)
        for name, reading, conf, margin, val, batch in (
                ("chaos_reading", bh.READING_CHAOS, args.chaos_reading_conf,
                 args.chaos_reading_margin, args.chaos_reading_validation,
                 args.chaos_reading_batch),
                ("class_reading", bh.READING_CLASS, args.chaos_conf, args.lle32_margin,
                 args.validation, args.batch)):
            try:
                bh.set_regime_reading(reading)
                bh.set_chaos_confidence(conf)
                bh.set_lle32_margin(margin)
                vms, vst = run(kv, args.seed_base, 1, val, batch)
                sc = scored(kv)
                variants[name] = {
                    "ms_per_step": vms, "speedup_vs_plain_fp64": plain_ms / vms, "steps": kv,
                    **breakdown(kv, vst),
                    "chaos_confidence": conf, "quad_confidence": args.quad_conf,
                    "lle32_margin": margin, "validation": val, "batch": batch,
                    "agreement": {k: summarise(np, sc, k)
                                  for k in ("regime", "chaos", "class", "count_exact")},
                    "lle_abs_err_mean": float(np.mean([a["lle_abs_err_mean"] for a in sc])),
                    "resolved_fraction": {
                        "full": float(np.mean([st["resolved_frac"] for st in vst])),
                        "peak_only": float(np.mean([st["peak_only"] for st in vst])) / cells,
                        "lle_only": float(np.mean([st["lle_only"] for st in vst])) / cells}}
            finally:
                bh.set_regime_reading(bh.READING_CLASS_AND_CHAOS)
                bh.set_chaos_confidence(args.chaos_conf)
                bh.set_lle32_margin(args.lle32_margin)
        variants["chaos_reading"]["reading"] = ("regime = chaos bit (LLE > 0.01) only; peak "
                                                "counts predicted, their agreement reported")
        variants["chaos_reading"]["gate_passed"] = \
            variants["chaos_reading"]["agreement"]["chaos"]["mean"] >= 0.99
        variants["class_reading"]["gate_passed"] = \
            variants["class_reading"]["agreement"]["class"]["mean"] >= 0.99
        variants["class_reading"]["reading"] = ("regime = peak-count class only; LLE values "
                                                "predicted, their error reported")

    gate = agree["mean"] >= 0.99
    out.update(value=cells / (ms / 1000), unit="grid points/s", ms_per_step=ms,
               config={"workload": f"surrogate-refined 2D LLE + DBSCAN peak-count map, "
                                   f"Rossler {n}x{n} (coarse {args.coarse}x{args.coarse}, "
                                   f"20 epochs; class confidence < {args.conf} or class "
                                   f"boundary: peak count re-solved in FP64; chaos "
                                   f"confidence < {args.chaos_conf}, regime boundary, or a "
                                   f"coarse quad with mixed or mispredicted exact chaos "
                                   f"bits (chaos confidence < {args.quad_conf}): "
                                   f"LLE re-solved, FP32 bulk where it lies >= "
                                   f"{args.lle32_margin} from the threshold and FP64 "
                                   f"nearer, with the FP64 peak count when the class is "
                                   f"flagged too or the regime changed)",
                       "grid": n, "coarse": args.coarse, "ranges": RANGES, "h": H,
                       "peak": PEAK, "lle": LLE, "chaos_threshold": CHAOS,
                       "regime": "(clamp(peak count, 1, 16), LLE > 0.01)",
                       "seeds": [args.seed_base, args.seed_base + K - 1]})
    if args.lle32_margin >= 0:
        out["dtype"] = "f64+f32"
    out["valid"] = gate
    out["agreement_gate"] = {"threshold": 0.99, "statistic": "mean over the timed steps",
                             "passed": gate}
    out["surrogate"] = {
        "regime_agreement": agree["mean"],
        "regime_agreement_per_step": agree,
        "chaos_agreement": summarise(np, per, "chaos"),
        "class_agreement": summarise(np, per, "class"),
        "count_exact_agreement": summarise(np, per, "count_exact"),
        "lle_abs_err_mean": float(np.mean([a["lle_abs_err_mean"] for a in per])),
        "batch": args.batch,
        "plain_fp64_ms": plain_ms,
        "plain_fp64_ms_runs": plain,
        "plain_fp64_value": cells / (plain_ms / 1000),
        "speedup_vs_plain_fp64": plain_ms / ms,
        "resolved_fraction": last["resolved_frac"],
        "peak_only_fraction": last["peak_only"] / cells,
        "lle_only_fraction": last["lle_only"] / cells,
        "lle_only_recounted_fraction": last["lle_only_recounted"] / cells,
        "lle_only_fp32_fraction": last["lle_only_fp32"] / cells,
        "lle32_margin": args.lle32_margin,
        "fp32_lle_cells": fp32_lle,
        "surrogate_on_lle_only": {"regime_acc": last["lle_only_regime_acc"]},
        "surrogate_on_peak_only": {"class_acc": last["peak_only_class_acc"]},
        "partial_resolve": not args.full_resolve,
        "phase_ms": {k: last[k] for k in ("ms_coarse", "ms_train", "ms_infer", "ms_resolve",
                                          "ms_flag", "ms_solve", "ms_score")},
        "network_share_of_step": fixed_ms / ms,
        "surrogate_on_flagged": {"class_acc": last["flagged_class_acc"],
                                 "regime_acc": last["flagged_regime_acc"]},
        "surrogate_on_validation_sample": {"cells": last["validation"],
                                           "class_acc": last["valid_class_acc"],
                                           "regime_acc": last["valid_regime_acc"]},
        "train_loss": {"ce": last["loss_ce"], "mse": last["loss_mse"]},
        "noise_ceiling": noise_ceiling(np, ref, n, plain_ms, fixed_ms, cost),
        "variants": variants,
    }
    out["kernels"] = rf
    dom = max((k for k in rf if k != "k_dbscan_wave"), key=lambda k: rf[k]["ms_per_step"])
    d = rf[dom]
    ratio = d.get("traffic_ratio")
    out["roofline"] = {"bound": d["bound"], "kernel": dom, "achieved": d["achieved"],
                       "peak": d["peak"], "unit": d["unit"], "frac": d["frac"],
                       "traffic": d.get("traffic"), "traffic_ratio": ratio,
                       "avg_launch_ms": d["avg_launch_ms"], "launches": d["launches"],
                       "flop_per_launch": d["work_per_launch"],
                       "note": "the kernel with the most time in the timed steps (every kernel "
                               "in 'kernels'); FP64 vector ALU, no-contraction FP64 reaches at "
                               "most 50% of the FMA-counted peak; traffic = HBM bytes written "
                               "per launch (counter ratio x algorithmic peak-row bytes)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="surrogate",
                    choices=["surrogate", "plain", "lle", "tiles", "dbscan", "kde", "bif1d", "drivers",
                             "copy"])
    ap.add_argument("--n", type=int, default=None, help="fine lattice (4096 / 1024)")
    ap.add_argument("--coarse", type=int, default=256)
    ap.add_argument("--nt", type=float, default=LLE["NT"],
                    help="lle workload: renormalisation interval in time units (0.5 in the "
                         "configuration; larger values isolate the renormalisations' cost)")
    ap.add_argument("--conf", type=float, default=0.95, help="class softmax confidence threshold")
    ap.add_argument("--chaos-conf", type=float, default=0.95,
                    help="chaos-logit confidence threshold (chosen on held-out seeds 1000-1002)")
    ap.add_argument("--quad-conf", type=float, default=2.0,
                    help="coarse-quad chaos rule: cells in coarse quads with mixed exact chaos "
                         "bits, or whose corners the network gets wrong, re-solve their LLE "
                         "below this chaos confidence (0 off, above 1 always)")
    ap.add_argument("--lle32-margin", type=float, default=0.02,
                    help="regime-only re-solves: FP32 bulk LLE at least this far from the "
                         "chaos threshold, FP64 nearer (negative: all FP64)")
    ap.add_argument("--validation", type=float, default=0.002)
    ap.add_argument("--seed-base", type=int, default=0,
                    help="surrogate seed of the first timed step (held-out sweeps use others)")
    ap.add_argument("--plain-steps", type=int, default=3,
                    help="timed runs of the plain FP64 map (after a full-size warm-up)")
    ap.add_argument("--variant-steps", type=int, default=3,
                    help="timed steps of each variant (all-FP64 re-solves, chaos-only and "
                         "class-only readings); 0 skips them")
    ap.add_argument("--chaos-reading-conf", type=float, default=0.6,
                    help="chaos-only reading variant: chaos-confidence threshold")
    ap.add_argument("--chaos-reading-margin", type=float, default=0.01,
                    help="chaos-only reading variant: FP32 LLE margin")
    ap.add_argument("--chaos-reading-batch", type=int, default=4096,
                    help="chaos-only reading variant: training batch rows")
    ap.add_argument("--chaos-reading-validation", type=float, default=0.0,
                    help="chaos-only reading variant: validation fraction (its agreement is "
                         "measured here against the plain map)")
    ap.add_argument("--batch", type=int, default=2048,
                    help="training batch rows (20 epochs of the 256^2 coarse map)")
    ap.add_argument("--cpu-cells", type=int, default=32768,
                    help="cells of the CPU baseline sample (whole lattice rows, ~6 s per run on "
                         "16 cores, median of 3 runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="surrogate workload: skip the brief config-3 / config-4 summaries")
    ap.add_argument("--fp32", action="store_true",
                    help="dbscan workload: also time the FP32-trajectory map and report its "
                         "agreement with the FP64 map")
    ap.add_argument("--full-resolve", action="store_true",
                    help="re-solve every flagged cell in full (no peak-count-only re-solves)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch tile and training work directly instead of replaying graphs")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import bifhip as bh

    bh.load()
    if bh.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    bh.reset_context()
    bh.set_perturbation(bh.PERT_AXIS_X)
    bh.set_graph(not args.no_graph)
    bh.set_partial_resolve(not args.full_resolve)
    bh.set_chaos_confidence(args.chaos_conf)
    bh.set_quad_confidence(args.quad_conf)
    bh.set_lle32_margin(args.lle32_margin)
    n = args.n or (4096 if args.workload == "surrogate" else 1024)
    cells = n * n
    d_peak = bh.DeviceArray((cells,), np.int32)
    d_lle = bh.DeviceArray((cells,), np.float64)
    out = {"metric": METRIC, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (a, c) parameter lattice of the Rossler system"}

    extra = {"tiles": run_tiles, "copy": run_copy, "lle": run_lle, "dbscan": run_dbscan,
             "kde": run_kde, "bif1d": run_bif1d, "drivers": run_drivers}
    if args.workload in extra:
        extra[args.workload](bh, np, args, out)
        print(json.dumps(out), flush=True)
        return
    if args.workload == "plain":
        for _ in range(args.warmup):
            plain_map(bh, n, d_peak, d_lle)
        bh.synchronize()
        bh.set_timing(True)
        bh.reset_kernel_times()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plain_map(bh, n, d_peak, d_lle)
        bh.synchronize()
        dt = time.perf_counter() - t0
        ms = dt * 1000 / args.steps
        out.update(value=cells / (ms / 1000), unit="grid points/s", ms_per_step=ms,
                   config={"workload": f"plain FP64 2D LLE + DBSCAN peak-count map, Rossler {n}x{n}",
                           "grid": n, "ranges": RANGES, "h": H, "peak": PEAK, "lle": LLE})
        out["roofline"] = roofline(bh, cells * args.steps)
        out["kernels_ms_per_step"] = {k: bh.kernel_time(k)[0] / args.steps
                                      for k in ("lle_peaks", "dbscan")}
    else:
        run_surrogate(bh, np, args, out, n, d_peak, d_lle)
        if not args.no_configs:
            bh.set_timing(False)
            c4 = run_config4(bh, np)
            out["configs"] = {
                "lle_fp32_bulk_1024": run_config3(bh, np),
                "ls_uneven_tiles_4096x16x16": {k: c4[k] for k in (
                    "ms_per_map", "autotuner", "fp64_tiles_ms", "fp32_vs_fp64")},
                "note": "configs 3 and 4 in brief, one map each (their own workloads: "
                        "--workload lle / tiles)"}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(n, max(1, args.cpu_cells // n))
    print(json.dumps(out), flush=True)
    if out.get("valid") is False:
        print("bench.py: mean regime agreement below 0.99: the line is marked invalid",
              file=sys.stderr)
        sys.exit(3)


if __name__ == "__main__":
    main()
