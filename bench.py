#!/usr/bin/env python3
"""Headline benchmark: grid points/s of the 2D LLE + peak-count map.

One step = one whole map over the synthetic (a, c) parameter lattice of
the Rossler system, inputs resident on the GPU, outputs left in HBM:
  * peak count: DBSCAN cluster count of (amplitude, interval) peak pairs,
    transient 1000, record 500 time units, h = 0.01, eps = 0.01;
  * LLE: Benettin two-trajectory, transient 500, measure 1000 time units,
    renormalisation every 0.5, perturbation 1e-8 along x.
Workloads:
  plain    plain FP64 path (reference operation order) on an n x n grid
Prints one JSON line (rank 0).  --gpus above 1 is refused: this build
targets one GPU.
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
FP64_VALU_PEAK = 78.6e12     # MI355X FP64 vector, FMA counted as 2
STEP_FLOP = 12               # Rossler sequential Euler step, no contraction
RENORM_FLOP = 20


def lle_flop_per_cell():
    skip = int(LLE["transient"] / H)
    nNT = int(LLE["NT"] / H)
    nb = int(LLE["tMax"] / LLE["NT"])
    return STEP_FLOP * (skip + 2 * nb * nNT) + RENORM_FLOP * nb


def peak_flop_per_cell():
    return STEP_FLOP * (int(PEAK["transient"] / H) + int(PEAK["tMax"] / H))


def run_plain(bh, n, d_peak, d_lle):
    N = int(PEAK["tMax"] / H)
    bh.dbscan_map_device(d_peak.ptr, PEAK["tMax"], n, H, IC, RANGES, [0, 2], 0, 0.0, N,
                         PEAK["transient"], VALS, 1, PEAK["eps"])
    bh.lle_map_device(d_lle.ptr, 2, LLE["tMax"], LLE["NT"], n, H, LLE["eps"], IC, RANGES,
                      [0, 2], 0.0, LLE["transient"], VALS)


def cpu_baseline(n, rows):
    """Oracle (the C restatement of the reference's per-trajectory code,
    OpenMP over cells) on `rows` whole rows of the same map."""
    import orclib as O
    O.lib()
    N = int(PEAK["tMax"] / H)
    step = max(1, n // rows)
    cells = 0
    t0 = time.perf_counter()
    for r in range(0, n, step)[:rows]:
        O.dbscan_cells(PEAK["tMax"], n, H, IC, RANGES, [0, 2], 0, 0.0, N, PEAK["transient"], VALS,
                       1, PEAK["eps"], r * n, n)
        O.lle_cells(2, LLE["tMax"], LLE["NT"], n, H, LLE["eps"], IC, RANGES, [0, 2], 0.0,
                    LLE["transient"], VALS, O.PERT_AXIS_X, 0, r * n, n)
        cells += n
    dt = time.perf_counter() - t0
    return {"value": cells / dt, "unit": "grid points/s", "cores": O.lib().orc_threads(),
            "kind": "port",
            "sample": f"{cells} cells ({rows} rows of the {n}x{n} map), DBSCAN peak count + LLE, "
                      f"{dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="plain", choices=["plain"])
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--cpu-rows", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
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
    n = args.n
    d_peak = bh.DeviceArray((n * n,), np.int32)
    d_lle = bh.DeviceArray((n * n,), np.float64)

    for _ in range(args.warmup):
        run_plain(bh, n, d_peak, d_lle)
    bh.synchronize()
    bh.set_timing(True)
    bh.reset_kernel_times()
    bh.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run_plain(bh, n, d_peak, d_lle)
    bh.synchronize()
    dt = time.perf_counter() - t0
    bh.set_timing(False)

    ms_step = dt * 1000 / args.steps
    cells = n * n
    lle_ms, lle_n = bh.kernel_time("lle")
    pk_ms, pk_n = bh.kernel_time("bif_peaks")
    db_ms, db_n = bh.kernel_time("dbscan")
    avg_lle = lle_ms / max(lle_n, 1)
    flop_launch = lle_flop_per_cell() * cells  # one launch covers the whole map
    achieved = flop_launch / (avg_lle / 1000) / 1e12
    out = {
        "metric": "grid points/s, 2D LLE+peak map; surrogate-refined vs plain FP64 @>=99% agree",
        "value": cells / (ms_step / 1000),
        "unit": "grid points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (a,c) parameter lattice of the Rossler system",
        "config": {"workload": f"plain FP64 2D LLE + DBSCAN peak-count map, Rossler {n}x{n}",
                   "grid": n, "a_range": RANGES[:2], "c_range": RANGES[2:], "h": H,
                   "peak": PEAK, "lle": LLE, "parallelism": "single GPU"},
        "roofline": {"bound": "valu_fp64", "kernel": "k_lle64", "achieved": achieved,
                     "peak": FP64_VALU_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / (FP64_VALU_PEAK / 1e12), "traffic": None,
                     "flop_per_launch": flop_launch, "avg_launch_ms": avg_lle},
        "kernels_ms_per_step": {"lle": lle_ms / args.steps, "bif_peaks": pk_ms / args.steps,
                                "dbscan": db_ms / args.steps},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(n, args.cpu_rows)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
