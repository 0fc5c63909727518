"""Benchmark: kNN-neighborhood features on one MI355X.

Workload (the plain GPU path vs CPU configuration): 4,000,000 seeded
synthetic ALS points (500 m x 500 m, 16 pts/m^2, 70% terrain / 30% clusters,
0.05 m noise, 0.01 m LAS quantum), filters.normal knn=16 chained with
filters.covariancefeatures knn=16 (Dimensionality) on one resident buffer.
A step = one pass of the hot path over the resident cloud: grid build, kNN
search + covariance, eigensolve + 8 output columns, exact re-solve of tie
rows; inputs and outputs stay in HBM.

Prints ONE JSON line (rank 0).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--points", type=int, default=4_000_000)
    ap.add_argument("--knn", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(tiles shard across GPUs without collectives; see the design notes)"}))
        sys.exit(2)

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

    for _ in range(args.warmup):
        cloud.invalidate()
        cloud.run_device(stages, cols, status)
    ctx.sync()
    per = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cloud.invalidate()  # every step rebuilds the index: a new cloud per step
        cloud.run_device(stages, cols, status)
        per.append(ctx.stats())
    ctx.sync()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    value = n / (ms / 1e3)

    st = per[-1]
    ms_search = float(np.mean([p["ms_search"] for p in per]))
    ms_index = float(np.mean([p["ms_index"] for p in per]))
    ms_feat = float(np.mean([p["ms_features"] for p in per]))
    # Algorithmic bytes of the search kernel per point: own FP64 XYZ (24 B),
    # the FP64 covariance it writes (48 B) and the row flag (1 B).
    alg_bytes = 73.0 * n
    achieved = alg_bytes / (ms_search / 1e3) / 1e9
    out = {
        "metric": "points/s kNN+covariance features (normal+covariancefeatures knn=16, 4M pts)",
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "C2: normal+covariancefeatures knn=16, 4M-point seeded ALS tile",
                   "points": n, "knn": knn},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8000.0, "traffic": None,
                     "kernel": "k_knn (search + covariance)"},
        "phases_ms": {"index": ms_index, "search": ms_search, "features": ms_feat},
        "tie_rows": st["tie_rows"],
        "grid": {"cell": st["cell"], "cell_z": st["cell_z"], "gx": st["grid_x"],
                 "gy": st["grid_y"], "gz": st["grid_z"], "key_bits": st["key_bits"]},
    }
    print(json.dumps(out))
    for c in cols:
        ctx.free(c)
    for s in status:
        ctx.free(s)


if __name__ == "__main__":
    main()
