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
sample of rows, single-threaded (PDAL's default), extrapolated to the cloud;
cpu_baseline_parallel repeats the stages on up to 16 host threads.

More legs run by default and are reported beside the headline value:
  batched  -- 256 seeded tiles (sizes log-uniform in [10k, 2M], knn in
              {8, 16, 32}, density 2-50 pts/m^2) as one batch vs the same
              tiles one at a time through the plain path;
  serving  -- Poisson arrivals of tile requests (sizes log-uniform in
              [5k, 4M], knn in {10, 16}) into the serving loop: p50/p99
              latency (submit -> result on the host) and sustained points/s;
  precision -- the precision-stress cloud (20M points, 5% planar roofs, 2%
              collinear wires, eigenvalues + covariancefeatures knn=32): per
              column max and p99 error against the extended-precision truth
              on a random sample of rows.

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
# xyz (24 B) and grid key (4 B) read once, its covariance record written (one
# 64-byte record: six FP64 entries, the row-flags word and one pad word).
SEARCH_BYTES_PER_POINT = 92.0
# HBM traffic of one k_knn launch on the bench workload (4,000,000 points,
# knn=16), from rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE runs of this
# script (KiB per dispatch), each scaled by the same counter on a 1 GiB copy
# kernel in the same runs: FETCH_SIZE reads 0.5 of the copied bytes there
# (factor 2.0), WRITE_SIZE 1.0 (factor 1.0).  kernel_commit names the commit
# whose k_knn source the counters were taken on; the runs, commands and raw
# counter values are in profiles/k_knn_plain_4M_knn16_traffic.md.
MEASURED_TRAFFIC = {"points": 4_000_000, "knn": 16, "fetch_kib": 83510.1,
                    "write_kib": 312652.9, "fetch_factor": 2.0, "write_factor": 1.0,
                    "kernel_commit": "6472501"}
# FP64 flops per candidate distance: 3 sub, 3 mul, 2 add.
FLOPS_PER_CANDIDATE = 8.0
HBM_PEAK_GBS = 8000.0
FP64_VECTOR_PEAK_TFLOPS = 78.6


def new_context(args):
    """A library context on GPU 0 with the --tune settings applied."""
    import gpupdal_amd as g
    ctx = g.Context(0)
    for kv in args.tune or []:
        key, val = kv.split("=", 1)
        ctx.set_tuning(key, float(val))
    return ctx


def gpu_leg(args):
    import numpy as np
    import gpupdal_amd as g

    n, knn = args.points, args.knn
    x, y, z = g.synth_cloud(n, tile=0, side=500.0)
    ctx = new_context(args)
    cloud = ctx.cloud(x, y, z)
    stages = [g.stage(g.STAGE_NORMAL, knn), g.stage(g.STAGE_COVFEAT, knn)]
    ncols = sum(g.stage_columns(s) for s in stages)
    cols = [ctx.alloc(n * 8) for _ in range(ncols)]
    status = [ctx.alloc(n) for _ in stages]

    # Timed loop.  Per-phase device times come from events recorded on the
    # library's streams and read after each call's final synchronisation,
    # so they add no wait between kernels.
    ctx.set_tuning("timing", 1)
    ctx.set_tuning("kd_overlap_min", args.kd_overlap_min)
    for _ in range(args.warmup):
        cloud.invalidate()
        cloud.run_device(stages, cols, status)
    ctx.sync()
    per = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cloud.invalidate()  # a new cloud every step: the index is rebuilt
        cloud.run_device(stages, cols, status)
        per.append(ctx.stats())
    ctx.sync()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps

    serial = None
    if args.serial:
        # The same step with the exact tree built after the search instead
        # of beside it.
        ctx.set_tuning("kd_overlap_min", 0)
        ctx.sync()
        s0 = time.perf_counter()
        for _ in range(3):
            cloud.invalidate()
            cloud.run_device(stages, cols, status)
        ctx.sync()
        serial = (time.perf_counter() - s0) * 1e3 / 3
        ctx.set_tuning("kd_overlap_min", args.kd_overlap_min)

    # Search counters in a separate pass.
    ctx.set_tuning("collect_stats", 1)
    cloud.invalidate()
    cloud.run_device(stages, cols, status)
    counters = ctx.stats()
    ctx.set_tuning("collect_stats", 0)

    def mean(key):
        return float(np.mean([p[key] for p in per]))

    ms_search = mean("ms_search")
    traffic = None
    mt = MEASURED_TRAFFIC
    if n == mt["points"] and knn == mt["knn"]:
        traffic = 1024.0 * (mt["fetch_kib"] * mt["fetch_factor"] +
                            mt["write_kib"] * mt["write_factor"])
    achieved = SEARCH_BYTES_PER_POINT * n / (ms_search / 1e3) / 1e9
    cand_per_s = counters["candidates"] / (ms_search / 1e3)
    fp64 = cand_per_s * FLOPS_PER_CANDIDATE / 1e12
    st = per[-1]
    out = dict(
        value=n / (ms / 1e3), ms_per_step=ms,
        roofline={"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                  "traffic_source": (None if traffic is None else
                                     "rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated; k_knn as of "
                                     "commit " + mt["kernel_commit"]),
                  "kernel": "k_knn (kNN search + FP64 covariance)",
                  "bytes_per_point": SEARCH_BYTES_PER_POINT,
                  "kernel_ms": ms_search},
        roofline_fp64={"bound": "fp64-valu", "achieved": fp64, "peak": FP64_VECTOR_PEAK_TFLOPS,
                       "unit": "TFLOP/s", "frac": fp64 / FP64_VECTOR_PEAK_TFLOPS,
                       "candidates_per_point": counters["candidates"] / n,
                       "flops_per_candidate": FLOPS_PER_CANDIDATE},
        # features: the rows the search did not flag, right after the search;
        # exact_tail: from there to the end of the flagged rows' features (tree
        # not hidden, exact re-solve, flagged rows' features).
        phases_ms={"index": mean("ms_index"), "search": ms_search,
                   "features": mean("ms_features"), "exact_tail": mean("ms_exact"),
                   "tree_beside_search": mean("ms_tree"), "device_total": mean("ms_total")},
        rows={"tie": int(st["tie_rows"]), "incomplete": int(st["incomplete_rows"])},
        grid={"cell_m": st["cell"], "cell_z_m": st["cell_z"], "gx": st["grid_x"],
              "gy": st["grid_y"], "gz": st["grid_z"], "key_bits": st["key_bits"]},
    )
    if serial is not None:
        out["ms_per_step_tree_after_search"] = serial
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


def gen_tiles(specs, workers=16):
    """Seeded tiles [(x, y, z)] for specs [(n, tile_seed_index, side)],
    generated in parallel threads (the generator releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    import gpupdal_amd as g

    def one(spec):
        n, idx, side = spec
        return g.synth_cloud(n, tile=idx, side=side)
    with ThreadPoolExecutor(max_workers=min(workers, os.cpu_count() or 1)) as ex:
        return list(ex.map(one, specs))


def compare_results(pairs, info_mask, acc=None):
    """Result check of a batched or served run against the same inputs run
    alone: pairs of (columns, statuses) host arrays, tallied into acc.
    Statuses must be equal outside the search's informational bits (a row
    re-solved by the exact search in one run and not the other); columns
    equal up to 1e-12 relative (the two summation orders of such a row)."""
    import numpy as np
    if acc is None:
        acc = {"rows": 0, "status_mismatch_rows": 0, "nan_mismatch": 0,
               "values_above_1e-12_rel": 0, "max_rel_diff": 0.0, "values": 0, "identical": 0}
    for (ca, sa), (cb, sb) in pairs:
        if sa:
            acc["rows"] += len(sa[0])
        for a, b in zip(sa, sb):
            acc["status_mismatch_rows"] += int(np.count_nonzero((a & ~info_mask) !=
                                                                (b & ~info_mask)))
        for a, b in zip(ca, cb):
            same = a.view(np.uint64) == b.view(np.uint64)
            acc["values"] += len(a)
            acc["identical"] += int(np.count_nonzero(same))
            na, nb = np.isnan(a), np.isnan(b)
            acc["nan_mismatch"] += int(np.count_nonzero(na != nb))
            d = ~same & ~na & ~nb
            if d.any():
                rel = np.abs(a[d] - b[d]) / np.maximum(np.maximum(np.abs(a[d]), np.abs(b[d])),
                                                       1e-300)
                acc["max_rel_diff"] = max(acc["max_rel_diff"], float(rel.max()))
                acc["values_above_1e-12_rel"] += int(np.count_nonzero(rel > 1e-12))
    acc["bit_identical_frac"] = acc["identical"] / max(acc["values"], 1)
    acc["ok"] = (acc["status_mismatch_rows"] == 0 and acc["nan_mismatch"] == 0 and
                 acc["values_above_1e-12_rel"] == 0)
    return acc


def c4_specs(ntiles, seed=0xC4):
    """The batched configuration: tile sizes log-uniform in [10k, 2M], knn in
    {8, 16, 32}, density uniform in [2, 50] points/m^2 (side = sqrt(n/density))."""
    import numpy as np
    rng = np.random.default_rng(seed)
    specs, knns = [], []
    for i in range(ntiles):
        n = int(np.exp(rng.uniform(np.log(10_000), np.log(2_000_000))))
        knn = int(rng.choice([8, 16, 32]))
        dens = rng.uniform(2.0, 50.0)
        specs.append((n, 1000 + i, float(np.sqrt(n / dens))))
        knns.append(knn)
    return specs, knns


def batched_leg(args):
    """Batched vs plain: the same tiles as one batch (one index build, one
    search, one exact re-solve, one feature pass) and one tile at a time
    through the plain path with a synchronisation per tile.  Inputs and
    outputs resident on the device; every step rebuilds every index.

    Each path runs with the context's default settings and with the table
    the autotuner builds (every (density, knn) bucket the tiles fall in,
    tuned on one tile of at least 100k points: column factor and search
    shape), and keeps the faster -- the autotuner's rule of keeping only
    what a whole-run timing shows faster, applied to the batch as a whole."""
    import numpy as np
    import gpupdal_amd as g
    specs, knns = c4_specs(args.tiles)
    tiles = gen_tiles(specs)
    n_total = sum(len(t[0]) for t in tiles)
    stages = [g.stage(g.STAGE_NORMAL, 16), g.stage(g.STAGE_COVFEAT, 16)]
    ncols = 8
    offs = np.cumsum([0] + [len(t[0]) for t in tiles])

    def autotune(ctx):
        tuned = set()
        for (x, y, z), k in zip(tiles, knns):
            if len(x) < 100_000:
                continue
            rho = len(x) / max((x.max() - x.min()) * (y.max() - y.min()), 1e-9)
            key = (int(np.floor(np.log2(rho))), int(np.ceil(np.log2(k + 1))))
            if key in tuned:
                continue
            tuned.add(key)
            cl = ctx.cloud(x, y, z)
            ctx.autotune(cl, [g.stage(g.STAGE_NORMAL, k), g.stage(g.STAGE_COVFEAT, k)],
                         factors=(0.45, 0.6, 0.75, 0.9))
            cl.close()
        return len(tuned)

    def measure(ctx, check=False):
        cols = [ctx.alloc(n_total * 8) for _ in range(ncols)]
        stat = [ctx.alloc(n_total) for _ in range(2)]
        # The plain path writes into its own columns when its results are
        # checked against the batch's.
        pcols = [ctx.alloc(n_total * 8) for _ in range(ncols)] if check else cols
        pstat = [ctx.alloc(n_total) for _ in range(2)] if check else stat
        b = ctx.batch(tiles)
        # Two untimed runs: the first runs launch by launch, the second
        # captures the step's graph; the timed steps replay it.
        for _ in range(2):
            b.invalidate()
            b.run_device(stages, cols, stat, tile_knn=knns)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.batch_steps):
            b.invalidate()
            b.run_device(stages, cols, stat, tile_knn=knns)
        ctx.sync()
        ms_b = (time.perf_counter() - t0) * 1e3 / args.batch_steps
        graph_mode = int(ctx.stats()["graph"])
        b.close()
        # Plain: one resident cloud per tile, one gf_run per tile (each
        # returns after its own synchronisation), outputs into the same
        # device columns.
        clouds = [ctx.cloud(*t) for t in tiles]

        def plain_step():
            for i, cl in enumerate(clouds):
                cl.invalidate()
                o = int(offs[i])
                st = [g.stage(g.STAGE_NORMAL, knns[i]), g.stage(g.STAGE_COVFEAT, knns[i])]
                cl.run_device(st, [c + 8 * o for c in pcols], [s + o for s in pstat])
        plain_step()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.batch_steps):
            plain_step()
        ctx.sync()
        ms_p = (time.perf_counter() - t0) * 1e3 / args.batch_steps
        for cl in clouds:
            cl.close()
        checked = None
        if check:
            # Every tile of the batch against the same tile run alone, one
            # column at a time through host memory.
            info = np.uint8(g.ST_TIE | g.ST_INCOMPLETE)
            sb = [ctx.copy_to_host(np.empty(n_total, np.uint8), p) for p in stat]
            sp = [ctx.copy_to_host(np.empty(n_total, np.uint8), p) for p in pstat]
            acc = compare_results([(([], sb), ([], sp))], info)
            del sb, sp
            for c in range(ncols):
                a = ctx.copy_to_host(np.empty(n_total, np.float64), cols[c])
                b = ctx.copy_to_host(np.empty(n_total, np.float64), pcols[c])
                acc = compare_results([(([a], []), ([b], []))], info, acc)
                del a, b
            acc["what"] = "every tile's columns and statuses, batch vs the tile run alone"
            checked = acc
            for p in pcols + pstat:
                ctx.free(p)
        for p in cols + stat:
            ctx.free(p)
        return ms_b, ms_p, graph_mode, checked

    ctx = new_context(args)
    ctx.set_tuning("timing", 0)
    ms_b_def, ms_p_def, graph_def, check = measure(ctx, check=True)
    ctx.close()
    tuned = 0
    ms_b_tun = ms_p_tun = float("inf")
    graph_tun = 0
    if not args.no_autotune:
        ctx = new_context(args)
        ctx.set_tuning("timing", 0)
        tuned = autotune(ctx)
        ms_b_tun, ms_p_tun, graph_tun, _ = measure(ctx)
        ctx.close()
    ms_batched = min(ms_b_def, ms_b_tun)
    ms_plain = min(ms_p_def, ms_p_tun)
    graph_mode = graph_tun if ms_b_tun < ms_b_def else graph_def
    return {"tiles": len(tiles), "points": int(n_total),
            "workload": "normal + covariancefeatures (Dimensionality), per-tile knn in {8,16,32}, "
                        "tile sizes log-uniform [10k, 2M], density U[2, 50] pts/m^2",
            "value": n_total / (ms_batched / 1e3), "unit": "points/s",
            "ms_per_step": ms_batched,
            "plain_value": n_total / (ms_plain / 1e3), "plain_ms_per_step": ms_plain,
            "batched_over_plain": ms_plain / ms_batched,
            "batched_ms": {"default": ms_b_def, "autotuned": ms_b_tun if tuned else None},
            "plain_ms": {"default": ms_p_def, "autotuned": ms_p_tun if tuned else None},
            "autotuned_buckets": tuned, "graph_replayed": graph_mode == 1,
            "check": check}


def las_leg(args, xyz):
    """The LAS boundary: the bench cloud as an uncompressed LAS 1.2 file
    image (point format 1, 28-byte records) in host memory, decoded into a
    resident cloud.  Device decode time from HIP events; the wall time adds
    the copy of the records over PCIe."""
    import numpy as np
    import gpupdal_amd as g
    from gpupdal_amd.las import las_image
    x, y, z = xyz
    img, _ = las_image(x, y, z, offset=(float(np.floor(x.min())), float(np.floor(y.min())),
                                        0.0), point_format=1)
    ctx = new_context(args)
    for _ in range(2):
        ctx.cloud_from_las(img).close()
    dec, wall = [], []
    for _ in range(args.las_steps):
        t0 = time.perf_counter()
        cl = ctx.cloud_from_las(img)
        wall.append((time.perf_counter() - t0) * 1e3)
        dec.append(cl.las_info["ms_decode"])
        cl.close()
    ctx.close()
    n = len(x)
    L = 28
    ms_dec = sorted(dec)[len(dec) // 2]
    ms_wall = sorted(wall)[len(wall) // 2]
    achieved = (L + 24) * n / (ms_dec / 1e3) / 1e9
    return {"points": n, "point_format": 1, "record_bytes": L,
            "workload": "uncompressed LAS 1.2 image of the bench cloud in host memory -> "
                        "resident FP64 cloud",
            "ms_decode": ms_dec, "decode_points_per_s": n / (ms_dec / 1e3),
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "bytes_per_point": L + 24},
            "ms_with_upload": ms_wall, "points_per_s_with_upload": n / (ms_wall / 1e3)}


def optimal_leg(args, xyz):
    """filters.optimalneighborhood (min_k 10, max_k 14) on the bench cloud:
    index, ordered 14-neighbor lists with their exact re-solve, and the
    per-prefix eigenentropy; the index and tree are rebuilt every step and
    the two output columns copied to the host."""
    import gpupdal_amd as g
    ctx = new_context(args)
    cl = ctx.cloud(*xyz)
    cl.invalidate()
    cl.optimal_neighborhood(10, 14)
    ms = []
    for _ in range(args.las_steps):
        cl.invalidate()
        t0 = time.perf_counter()
        cl.optimal_neighborhood(10, 14)
        ms.append((time.perf_counter() - t0) * 1e3)
    cl.close()
    ctx.close()
    med = sorted(ms)[len(ms) // 2]
    n = len(xyz[0])
    return {"points": n, "min_k": 10, "max_k": 14, "ms_per_step": med,
            "points_per_s": n / (med / 1e3),
            "workload": "filters.optimalneighborhood on the bench cloud, index and tree "
                        "rebuilt each step, OptimalKNN/OptimalRadius copied to the host"}


def serving_leg(args):
    """Serving: Poisson arrivals of tile requests at a fixed rate (sizes
    log-uniform in [5k, 4M], knn in {10, 16}, normal + covariancefeatures)
    into the serving loop (two workers, 200 GB HBM budget); latency is
    submit -> result on the host."""
    import threading
    import numpy as np
    import gpupdal_amd as g
    rng = np.random.default_rng(0xC5)
    specs, knns = [], []
    for i in range(args.requests):
        n = int(np.exp(rng.uniform(np.log(5_000), np.log(4_000_000))))
        specs.append((n, 5000 + i, float(np.sqrt(n / 16.0))))
        knns.append(int(rng.choice([10, 16])))
    tiles = gen_tiles(specs)
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, size=len(tiles)))
    stages = [g.stage(g.STAGE_NORMAL, 10), g.stage(g.STAGE_COVFEAT, 10)]
    srv = g.Server(workers=2, hbm_budget_bytes=200 << 30)
    # Warm-up request (contexts, code objects, first allocations).
    t = srv.submit(*tiles[0], stages, knns[0])
    srv.result(t)
    tickets = [None] * len(tiles)

    def producer():
        t0 = time.perf_counter()
        for i, tile in enumerate(tiles):
            delay = arrivals[i] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            tickets[i] = srv.submit(*tile, stages, knns[i])
    th = threading.Thread(target=producer)
    th.start()
    th.join()

    kept = []

    def collect(tks, keep=False):
        lat, queue, service, first, last, batches = [], [], [], None, None, set()
        for tk in tks:
            cols_r, stat_r, (ts, tb, td, bid, bt) = srv.result(tk)
            if keep:
                kept.append(([c for cs in cols_r for c in cs], stat_r))
            lat.append(td - ts)
            queue.append(tb - ts)
            service.append(td - tb)
            first = ts if first is None else min(first, ts)
            last = td if last is None else max(last, td)
            batches.add(bid)
        return (np.array(lat) * 1e3, np.array(queue) * 1e3, np.array(service) * 1e3,
                last - first, len(batches))

    lat_ms, q_ms, s_ms, span_s, nbat = collect(tickets, keep=True)
    # Capacity: the same requests submitted all at once (the queue never
    # runs dry), the sustained rate from the first submit to the last result.
    sat = [srv.submit(*tiles[i], stages, knns[i]) for i in range(len(tiles))]
    _, _, _, sat_s, sat_bat = collect(sat)
    srv.close()
    # Every served request against the same tile run alone (host arrays).
    ctx = g.Context(0)
    pairs = []
    for i, tile in enumerate(tiles):
        cl = ctx.cloud(*tile)
        (cn, cf), (sn, sf) = cl.run([g.stage(g.STAGE_NORMAL, knns[i]),
                                     g.stage(g.STAGE_COVFEAT, knns[i])])
        cl.close()
        pairs.append((kept[i], (cn + cf, [sn, sf])))
    ctx.close()
    check = compare_results(pairs, np.uint8(g.ST_TIE | g.ST_INCOMPLETE))
    check["what"] = "every request's columns and statuses, served vs the tile run alone"
    del pairs, kept[:]
    pts = sum(len(t[0]) for t in tiles)
    pct = lambda a, q: float(np.percentile(a, q))
    return {"requests": len(tiles), "rate_per_s": args.rate, "points": int(pts),
            "workload": "Poisson arrivals, sizes log-uniform [5k, 4M], knn in {10, 16}, "
                        "normal + covariancefeatures (Dimensionality), 2 workers, "
                        "200 GB HBM budget, host arrays in and out",
            "p50_ms": pct(lat_ms, 50), "p99_ms": pct(lat_ms, 99), "max_ms": float(lat_ms.max()),
            "queue_ms": {"p50": pct(q_ms, 50), "p99": pct(q_ms, 99)},
            "batch_ms": {"p50": pct(s_ms, 50), "p99": pct(s_ms, 99),
                         "what": "batch start to results on the host: uploads, device, downloads"},
            "sustained_points_per_s": pts / span_s, "batches": nbat,
            "offered_points_per_s": pts / float(arrivals[-1]),
            "capacity_points_per_s": pts / sat_s, "capacity_batches": sat_bat,
            "check": check}


def precision_leg(args, sample_rows=0):
    """Precision stress: 20M points with 5% exactly planar roofs and 2%
    collinear wires, eigenvalues + covariancefeatures knn=32 chained; every
    column against the extended-precision truth computed on the host from
    the same KD3Index neighbor lists.  Per feature: max and p99 relative
    error over well-conditioned rows (eigengap > 1e-3, smallest eigenvalue
    resolved), max absolute error over degenerate rows.  sample_rows > 0:
    the truth on that many random rows (the default run), else on every
    row."""
    import numpy as np
    import gpupdal_amd as g
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    n, knn = args.precision_points, 32
    side = float(np.sqrt(n / 80.0))
    x, y, z = g.synth_cloud(n, tile=3, side=side, roof_frac=0.05, wire_frac=0.02)
    ctx = new_context(args)
    cl = ctx.cloud(x, y, z)
    stages = [g.stage(g.STAGE_EIGENVALUES, knn), g.stage(g.STAGE_COVFEAT, knn)]
    t0 = time.perf_counter()
    (ev, cf), (evs, cfs) = cl.run(stages, init=np.nan)
    ms = (time.perf_counter() - t0) * 1e3
    st = ctx.stats()
    cl.close()
    ctx.close()
    threads = min(16, os.cpu_count() or 1)
    print(f"precision: GPU run done ({ms:.0f} ms); host KD3 lists", file=sys.stderr, flush=True)
    orc = ol.Oracle()
    tree = orc.kd3(x, y, z)
    if sample_rows:
        rows = np.sort(np.random.default_rng(0x3C).choice(n, min(sample_rows, n),
                                                          replace=False))
        nbr = tree.neighbors_rows(rows, knn + 1, threads=threads)
        ev = [c[rows] for c in ev]
        cf = [c[rows] for c in cf]
        evs, cfs = evs[rows], cfs[rows]
    else:
        nbr = tree.neighbors_all(knn + 1, threads=threads)
    del tree
    report = {"points": n, "knn": knn, "roof_frac": 0.05, "wire_frac": 0.02,
              "rows_checked": int(nbr.shape[0]),
              "rows": "random sample" if sample_rows else "all",
              "gpu_ms_host_io": ms, "tie_rows": int(st["tie_rows"]),
              "incomplete_rows": int(st["incomplete_rows"]), "columns": {}}
    for stage_cols, stat, kind in ((ev, evs, ol.STAGE_EIGENVALUES), (cf, cfs, ol.STAGE_COVFEAT)):
        ost = ol.stage(kind, knn)
        print(f"precision: host truth for stage {kind}", file=sys.stderr, flush=True)
        tcols, tstat, gap, ratio = orc.truth_from_lists(x, y, z, nbr, ost, threads=threads,
                                                        with_ratio=True)
        ok = ((stat & 0x3C) == 0) & (tstat == 0)
        well = ok & (gap > 1e-3) & (ratio > 1e-10)
        deg = ok & (ratio <= 1e-10)
        names = (["Eigenvalue0", "Eigenvalue1", "Eigenvalue2"] if kind == ol.STAGE_EIGENVALUES
                 else [ol.FEATURE_NAMES[b] for b in range(4)])
        for name, c, t in zip(names, stage_cols, tcols):
            rel = np.abs(c[well] - t[well]) / np.maximum(np.abs(t[well]), 1e-6)
            ab = np.abs(c[deg] - t[deg])
            report["columns"][name] = {
                "well_rows": int(well.sum()), "max_rel": float(rel.max()) if rel.size else 0.0,
                "p99_rel": float(np.percentile(rel, 99)) if rel.size else 0.0,
                "degenerate_rows": int(deg.sum()),
                "degenerate_max_abs": float(ab.max()) if ab.size else 0.0}
        report.setdefault("skipped_rows", {})[names[0][:5]] = int((~ok).sum())
    return report


def host_cpu():
    """The host's logical CPU count and model name (from /proc/cpuinfo)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model,
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
            else None}


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
    host = host_cpu()
    out = {"value": n / total, "unit": "points/s", "cores": 1, "kind": "port", **host,
           "sample": f"KD build over all {n} points ({t_build:.2f} s) + normal and "
                     f"covariancefeatures knn={args.knn} on rows [0, {rows}) "
                     f"({t_rows:.2f} s), extrapolated to {n} rows",
           "seconds_estimated": total}
    # The same stages with contiguous row chunks on several host threads (the
    # upstream filter's threads option); the KD build stays single-threaded,
    # as upstream's is per view.
    threads = max(1, min(16, os.cpu_count() or 1))
    prow = min(n, rows * 4)
    t0 = time.perf_counter()
    orc.stage_run_range(tree, ol.stage(ol.STAGE_NORMAL, args.knn), 0, prow, threads=threads)
    orc.stage_run_range(tree, ol.stage(ol.STAGE_COVFEAT, args.knn), 0, prow, threads=threads)
    t_prow = time.perf_counter() - t0
    ptotal = t_build + t_prow * n / prow
    parallel = {"value": n / ptotal, "unit": "points/s", "cores": threads, "kind": "port",
                **host,
                "sample": f"as the single-threaded sample, rows [0, {prow}) on {threads} "
                          f"threads ({t_prow:.2f} s), extrapolated to {n} rows",
                "seconds_estimated": ptotal}
    return out, parallel


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
    ap.add_argument("--tiles", type=int, default=256, help="tiles of the batched leg")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-vs-plain leg")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched leg: no autotuning of the tiles' buckets")
    ap.add_argument("--requests", type=int, default=64, help="requests of the serving leg")
    ap.add_argument("--rate", type=float, default=120.0, help="serving arrivals per second")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    ap.add_argument("--only-serving", action="store_true", help="run the serving leg alone")
    ap.add_argument("--no-las", action="store_true",
                    help="skip the LAS decode and optimal-neighborhood legs")
    ap.add_argument("--las-steps", type=int, default=7)
    ap.add_argument("--precision", action="store_true",
                    help="only the precision-stress report (20M points, roofs, wires)")
    ap.add_argument("--precision-points", type=int, default=20_000_000)
    ap.add_argument("--precision-sample", type=int, default=250_000,
                    help="rows of the precision leg checked against the truth in the default run")
    ap.add_argument("--no-precision", action="store_true", help="skip the precision leg")
    ap.add_argument("--kd-overlap-min", type=int, default=131072,
                    help="clouds of at least this many points build the exact tree beside "
                         "the search (0: after it)")
    ap.add_argument("--serial", action="store_true",
                    help="also time the step with the exact tree built after the search")
    ap.add_argument("--tune", action="append", metavar="KEY=VALUE",
                    help="a library tuning setting for every context (repeatable)")
    ap.add_argument("--only-batched", action="store_true",
                    help="only the batched-vs-plain leg (for profiling it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: --gpus must be 1. Tiles are "
                          "independent and would shard across GPUs without collectives."}))
        sys.exit(2)

    if args.precision:
        print(json.dumps({"precision": precision_leg(args)}))
        return
    if args.only_batched:
        print(json.dumps({"batched": batched_leg(args)}))
        return
    if args.only_serving:
        print(json.dumps({"serving": serving_leg(args)}))
        return
    gpu, xyz = gpu_leg(args)
    batched = None if args.no_batched else batched_leg(args)
    serving = None if args.no_serving else serving_leg(args)
    las = None if args.no_las else las_leg(args, xyz)
    optimal = None if args.no_las else optimal_leg(args, xyz)
    precision = (None if args.no_precision
                 else precision_leg(args, sample_rows=args.precision_sample))
    cpu, cpu_parallel = (None, None) if args.no_cpu else cpu_leg(args, xyz)
    line = {
        "metric": "points/s kNN+covariance features, batched vs plain GPU; p50/p99 tile latency",
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
        **({"ms_per_step_tree_after_search": gpu["ms_per_step_tree_after_search"]}
           if "ms_per_step_tree_after_search" in gpu else {}),
        "grid": gpu["grid"],
        "batched": batched,
        "serving": serving,
        "las_read": las,
        "optimal_neighborhood": optimal,
        "precision": precision,
        "cpu_baseline": cpu,
        "cpu_baseline_parallel": cpu_parallel,
    }
    if cpu:
        line["speedup_vs_cpu_baseline"] = gpu["value"] / cpu["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
