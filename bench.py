"""Benchmark: kNN-neighborhood features on one MI355X.

Workload (the plain GPU path vs CPU configuration): 4,000,000 seeded
synthetic ALS points (500 m x 500 m, 16 pts/m^2, 70% terrain / 30% clusters
and buildings, 0.05 m noise, 0.01 m LAS quantum), filters.normal knn=16
chained with filters.covariancefeatures knn=16 (Dimensionality) on one
resident buffer.  A step is one pass of the hot path over the resident cloud:
grid index build, kNN search + covariance, exact KD3 build and re-solve of
the rows with boundary distance ties, eigensolve + the 8 output columns.
Inputs and outputs stay in HBM; the index is rebuilt every step.

The CPU baseline is the oracle's restatement of PDAL's CPU path (KD3Index +
computeCovariance + the Eigen 3x3 solver + the stage formulas) timed on this
host: the KD build over the whole cloud plus both stages over a bounded
sample of rows, single-threaded (PDAL's default), extrapolated to the cloud.

Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic bytes per point of the search kernel (k_knn): the query's FP64
# xyz (24 B) read once, its FP64 covariance (48 B) and row flag (1 B) written.
SEARCH_BYTES_PER_POINT = 73.0
# FP64 flops per candidate distance: 3 sub, 3 mul, 2 add.
FLOPS_PER_CANDIDATE = 8.0
HBM_PEAK_GBS = 8000.0
FP64_VECTOR_PEAK_TFLOPS = 78.6


def gpu_leg(args):
    import numpy as np
    import gpupdal_amd as g

    n, knn = args.points, args.knn
    x, y, z = g.synth_cloud(n, tile=0, side=500.0)
    ctx = g.Context(0)
    cloud = ctx.cloud(x, y, z)
    stages = [g.stage(g.STAGE_NORMAL, knn), g.stage(g.STAGE_COVFEAT, knn)]
    ncols = sum(g.stage_columns(s) for s in stages)
    cols = [ctx.alloc(n * 8) for _ in range(ncols)]
    status = [ctx.alloc(n) for _ in stages]

    # Timed loop: no per-phase timers (they synchronize between kernels).
    ctx.set_tuning("timing", 0)
    for _ in range(args.warmup):
        cloud.invalidate()
        cloud.run_device(stages, cols, status)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cloud.invalidate()  # a new cloud every step: the index is rebuilt
        cloud.run_device(stages, cols, status)
    ctx.sync()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps

    # Phase breakdown and the dominant kernel's duration (events on the
    # library's stream), plus search counters, in separate passes.
    ctx.set_tuning("timing", 1)
    per = []
    for _ in range(3):
        cloud.invalidate()
        cloud.run_device(stages, cols, status)
        per.append(ctx.stats())
    ctx.set_tuning("collect_stats", 1)
    cloud.invalidate()
    cloud.run_device(stages, cols, status)
    counters = ctx.stats()
    ctx.set_tuning("collect_stats", 0)

    def mean(key):
        return float(np.mean([p[key] for p in per]))

    ms_search = mean("ms_search")
    achieved = SEARCH_BYTES_PER_POINT * n / (ms_search / 1e3) / 1e9
    cand_per_s = counters["candidates"] / (ms_search / 1e3)
    fp64 = cand_per_s * FLOPS_PER_CANDIDATE / 1e12
    st = per[-1]
    out = dict(
        value=n / (ms / 1e3), ms_per_step=ms,
        roofline={"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                  "kernel": "k_knn (kNN search + FP64 covariance)",
                  "bytes_per_point": SEARCH_BYTES_PER_POINT,
                  "kernel_ms": ms_search},
        roofline_fp64={"bound": "fp64-valu", "achieved": fp64, "peak": FP64_VECTOR_PEAK_TFLOPS,
                       "unit": "TFLOP/s", "frac": fp64 / FP64_VECTOR_PEAK_TFLOPS,
                       "candidates_per_point": counters["candidates"] / n,
                       "flops_per_candidate": FLOPS_PER_CANDIDATE},
        phases_ms={"index": mean("ms_index"), "search": ms_search,
                   "exact": mean("ms_exact"), "features": mean("ms_features")},
        rows={"tie": int(st["tie_rows"]), "incomplete": int(st["incomplete_rows"])},
        grid={"cell_m": st["cell"], "cell_z_m": st["cell_z"], "gx": st["grid_x"],
              "gy": st["grid_y"], "gz": st["grid_z"], "key_bits": st["key_bits"]},
    )
    if args.copy_calibration:
        # A plain copy of known size: FETCH_SIZE / WRITE_SIZE counters on
        # this kernel give the scale factor for the search kernel's counters.
        cnt = args.copy_calibration // 8
        a, b = ctx.alloc(cnt * 8), ctx.alloc(cnt * 8)
        for _ in range(3):
            ctx.copy(b, a, cnt)
        ctx.free(a)
        ctx.free(b)
    for c in cols:
        ctx.free(c)
    for s in status:
        ctx.free(s)
    cloud.close()
    ctx.close()
    return out, (x, y, z)


def cpu_leg(args, xyz):
    """PDAL's CPU path (the oracle's restatement), single-threaded, on a
    bounded sample of rows; the KD build covers the whole cloud."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    x, y, z = xyz
    n = len(x)
    orc = ol.Oracle()
    t0 = time.perf_counter()
    tree = orc.kd3(x, y, z)
    t_build = time.perf_counter() - t0
    rows = min(args.cpu_rows, n)
    t0 = time.perf_counter()
    orc.stage_run_range(tree, ol.stage(ol.STAGE_NORMAL, args.knn), 0, rows, threads=1)
    orc.stage_run_range(tree, ol.stage(ol.STAGE_COVFEAT, args.knn), 0, rows, threads=1)
    t_rows = time.perf_counter() - t0
    total = t_build + t_rows * n / rows
    return {"value": n / total, "unit": "points/s", "cores": 1, "kind": "port",
            "sample": f"KD build over all {n} points ({t_build:.2f} s) + normal and "
                      f"covariancefeatures knn={args.knn} on rows [0, {rows}) "
                      f"({t_rows:.2f} s), extrapolated to {n} rows",
            "seconds_estimated": total}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--points", type=int, default=4_000_000)
    ap.add_argument("--knn", type=int, default=16)
    ap.add_argument("--cpu-rows", type=int, default=200_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="also run a plain device copy of BYTES (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: --gpus must be 1. Tiles are "
                          "independent and would shard across GPUs without collectives."}))
        sys.exit(2)

    gpu, xyz = gpu_leg(args)
    cpu = None if args.no_cpu else cpu_leg(args, xyz)
    line = {
        "metric": "points/s kNN+covariance features (normal+covariancefeatures knn=16, "
                  "4M-point ALS tile, plain single-cloud GPU path)",
        "value": gpu["value"],
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": gpu["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded ALS generator, tile 0)",
        "config": {"workload": "plain GPU path: filters.normal knn=16 + "
                               "filters.covariancefeatures knn=16 (Dimensionality) chained "
                               "on one resident 4,000,000-point cloud",
                   "points": args.points, "knn": args.knn},
        "roofline": gpu["roofline"],
        "roofline_fp64": gpu["roofline_fp64"],
        "phases_ms": gpu["phases_ms"],
        "rows_resolved_exactly": gpu["rows"],
        "grid": gpu["grid"],
        "cpu_baseline": cpu,
    }
    if cpu:
        line["speedup_vs_cpu_baseline"] = gpu["value"] / cpu["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
