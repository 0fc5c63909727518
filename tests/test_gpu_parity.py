"""GPU parity: the HIP path (through the C ABI) against the oracle.

Neighbor sets are bit-exact against the KD3Index restatement; rows whose
search ends in a boundary distance tie are flagged (and checked to really be
tie rows).  Eigensystems from the device solver equal the restated Eigen
solver bit for bit.  Features are within 1e-5 relative of the
extended-precision truth on well-conditioned rows.
"""
import numpy as np
import pytest

import gpupdal_amd as g
import oracle_lib as ol
from helpers import small_cloud, set_equal_rows, rel_err, truth_errors

pytestmark = pytest.mark.gpu


def _host_keys(x, y, z, gp):
    """Each point's grid key from the grid's parameters, by the documented
    rule: cell coordinate floor((v - origin) / edge) clamped to the grid,
    column = 8x8 tile (row-major) * 64 + Morton interleave of the low three
    bits (x in the even bits), key = (column << zbits) | zcell."""
    def cell(v, o, inv, dim):
        f = np.floor((v - o) * inv)
        f = np.where(np.isnan(f) | (f < 0), 0, f)
        return np.minimum(f, dim - 1).astype(np.uint64)
    cx = cell(x, gp["ox"], gp["inv_c"], gp["gx"])
    cy = cell(y, gp["oy"], gp["inv_c"], gp["gy"])
    cz = cell(z, gp["oz"], gp["inv_cz"], gp["gz"])
    tile = (cy >> 3) * gp["tiles_x"] + (cx >> 3)
    mort = np.zeros_like(cx)
    for b in range(3):
        mort |= ((cx >> b) & 1) << (2 * b)
        mort |= ((cy >> b) & 1) << (2 * b + 1)
    col = tile * 64 + mort
    return (col << gp["bz"]) | cz, col


def test_grid_invariants(ctx):
    """The index against a host restatement of the grid rule: every point's
    key, the sorted order (a stable sort by key), and the column starts."""
    x, y, z = small_cloud(50000, seed=3)
    cl = ctx.cloud(x, y, z)
    perm, keys, colstart, bits = cl.grid(11)
    gp = cl.grid_params(11)
    n = len(x)
    assert bits == gp["key_bits"] and len(colstart) == gp["ncols"] + 1
    hkeys, hcol = _host_keys(x, y, z, gp)
    assert hkeys.max() < (1 << gp["key_bits"])
    # the sorted order is a stable sort of the host keys by point id
    order = np.argsort(hkeys, kind="stable")
    assert np.array_equal(order.astype(np.uint32), perm), "sorted order"
    assert np.array_equal(keys.astype(np.uint64), hkeys[order]), "sorted keys"
    # column starts: first sorted position of each column (n past the last)
    scol = hcol[order]
    want = np.searchsorted(scol, np.arange(gp["ncols"] + 1, dtype=np.uint64), side="left")
    assert np.array_equal(colstart.astype(np.int64), want.astype(np.int64)), "column starts"
    assert colstart[-1] == n


@pytest.mark.parametrize("n,k,seed", [(20000, 11, 1), (30000, 17, 2), (15000, 33, 4)])
def test_knn_sets_vs_kd3(ctx, oracle, n, k, seed):
    """Every row's neighbor set equals KD3Index's; rows re-solved by the exact
    tree search (ties, uncertified) also match its order."""
    x, y, z = small_cloud(n, seed=seed, dup_frac=0.02)
    cl = ctx.cloud(x, y, z)
    ids, st = cl.knn(k)
    tree = oracle.kd3(x, y, z)
    ref = tree.neighbors_all(k, threads=8)
    eq = set_equal_rows(ids, ref)
    assert np.all(eq), f"{np.count_nonzero(~eq)} rows differ from KD3Index"
    flagged = (st & (g.ST_TIE | g.ST_INCOMPLETE)) != 0
    assert np.array_equal(ids[flagged], ref[flagged]), "re-solved rows keep KD3 order"
    # every row flagged as a tie really has a boundary tie (brute force)
    tie_rows = np.nonzero(st & g.ST_TIE)[0]
    if len(tie_rows):
        bid, bd2 = oracle.brute_knn(x, y, z, k + 1, queries=tie_rows[:2000])
        assert np.all(bd2[:, k - 1] == bd2[:, k])


def _tree_nodes(t):
    inner = t["divfeat"] >= 0
    return {(int(a), int(b)): (int(f), lo, hi) for a, b, f, lo, hi in
            zip(t["left"][inner], t["right"][inner], t["divfeat"][inner],
                t["divlow"][inner], t["divhigh"][inner])}


def _oracle_nodes(e):
    ref = {}
    for i in range(len(e["child1"])):
        if e["child1"][i] >= 0:
            ref[(int(e["left"][i]), int(e["right"][i]))] = (
                int(e["divfeat"][i]), e["divlow"][i], e["divhigh"][i])
    return ref


@pytest.mark.parametrize("subtree,lds,small,wave", [(4096, 512, 0, 256), (200, 0, 0, 256),
                                                    (1000, 4, 0, 101), (4096, 512, 0, 1024),
                                                    (4096, 512, 1, 256), (2048, 512, 0, 256)],
                         ids=["default", "global-levels", "mixed", "wave-1024", "one-wave-blocks",
                              "half-size-blocks"])
def test_kdtree_matches_oracle(ctx, oracle, subtree, lds, small, wave):
    """The GPU-built tree is nanoflann's: same vind order and node splits,
    whichever way the build divides the work (level-synchronous big nodes,
    per-workgroup subtrees of 4096, 2048 or 1024 points with workgroup levels
    and one-wave subtrees, or one-wave workgroups of <= 512 points; LDS-staged
    or global extents)."""
    ctx.set_tuning("kd_subtree", subtree)
    ctx.set_tuning("kd_minmax_lds", lds)
    ctx.set_tuning("kd_small_blocks", small)
    ctx.set_tuning("kd_wave_node", wave)
    try:
        for seed, n, dup in [(21, 30000, 0.05), (22, 5000, 0.3), (23, 150, 0.0),
                             (24, 120000, 0.02)]:
            x, y, z = small_cloud(n, seed=seed, dup_frac=dup)
            t = ctx.cloud(x, y, z).kdtree()
            e = oracle.kd3(x, y, z).export()
            assert np.array_equal(t["vind"].astype(np.uint64), e["vind"]), (seed, n)
            assert _tree_nodes(t) == _oracle_nodes(e), (seed, n)
            # Ids a subtree reserved and did not use hold empty ranges.
            assert np.count_nonzero(t["right"] > t["left"]) == len(e["child1"])
    finally:
        ctx.set_tuning("kd_subtree", 2048)
        ctx.set_tuning("kd_minmax_lds", 128)
        ctx.set_tuning("kd_small_blocks", 0)
        ctx.set_tuning("kd_wave_node", 256)


def test_knn_lattice_ties_vs_kd3(ctx, oracle):
    """A regular lattice puts many points at equal distances: most rows have
    boundary ties and go through the exact tree search."""
    g1 = np.arange(30, dtype=np.float64)
    X, Y, Z = np.meshgrid(g1, g1, np.arange(6, dtype=np.float64), indexing="ij")
    x, y, z = X.ravel() * 0.5 + 500000.0, Y.ravel() * 0.5 + 4000000.0, Z.ravel() * 0.5
    cl = ctx.cloud(x, y, z)
    for k in (9, 17):
        ids, st = cl.knn(k)
        ref = oracle.kd3(x, y, z).neighbors_all(k, threads=8)
        assert np.count_nonzero(st & g.ST_TIE) > len(x) // 4
        assert np.all(set_equal_rows(ids, ref))


@pytest.mark.parametrize("k,stride", [(5, 2), (8, 3), (11, 2)])
def test_knn_stride_vs_kd3(ctx, oracle, k, stride):
    x, y, z = small_cloud(12000, seed=31, dup_frac=0.02)
    ids, st = ctx.cloud(x, y, z).knn(k, stride=stride)
    tree = oracle.kd3(x, y, z)
    ref = tree.neighbors_all(k, stride=stride, threads=8)
    assert np.all(set_equal_rows(ids, ref))


def test_knn_tie_free_rows_vs_bruteforce(ctx, oracle):
    x, y, z = small_cloud(8000, seed=9)
    cl = ctx.cloud(x, y, z)
    ids, st = cl.knn(17)
    bid, bd2 = oracle.brute_knn(x, y, z, 17)
    ok = (st & (g.ST_TIE | g.ST_INCOMPLETE)) == 0
    assert np.array_equal(ids[ok], bid[ok]), "ordered (d2, id) lists"


def test_device_eigen_bits(ctx, oracle):
    rng = np.random.default_rng(5)
    m = 4000
    covs = []
    for t in range(m):
        a = rng.normal(size=(3, 3))
        c = a @ a.T
        if t % 4 == 1:
            c[0, 2] = c[2, 0] = 0
        if t % 4 == 2:
            c = np.diag(rng.uniform(0, 1, 3))
        if t % 5 == 3:
            c *= 1e-9
        covs.append(c)
    covs = np.array(covs)
    info, ev, vec = ctx.debug_eigen(covs)
    for t in range(m):
        i2, e2, v2 = oracle.eigen_sym3(covs[t])
        assert info[t] == i2
        assert np.array_equal(ev[t], e2), t
        assert np.array_equal(vec[t], v2), t


@pytest.mark.parametrize("n,knn", [(40000, 10), (30000, 16)])
def test_covfeat_all_vs_truth(ctx, oracle, n, knn):
    x, y, z = small_cloud(n, seed=11)
    st = g.stage(g.STAGE_COVFEAT, knn, features=g.F_ALL10)
    (cols,), (status,) = ctx.cloud(x, y, z).run([st])
    tree = oracle.kd3(x, y, z)
    nbr = tree.neighbors_all(knn + 1, threads=8)
    ost = ol.stage(ol.STAGE_COVFEAT, knn, features=ol.F_ALL10)
    tcols, tstat, gap, ratio = oracle.truth_from_lists(x, y, z, nbr, ost, with_ratio=True)
    ok = ((status & 0x3C) == 0) & (tstat == 0)
    assert np.count_nonzero(ok) > 0.95 * n
    errs = truth_errors(ost, cols, tcols, ok & (gap > 1e-3), ratio)
    assert max(errs.values()) <= 1.0, errs


def test_normal_eigen_chain_vs_pdal_oracle(ctx, oracle):
    x, y, z = small_cloud(30000, seed=12)
    stages = [g.stage(g.STAGE_NORMAL, 16), g.stage(g.STAGE_COVFEAT, 16)]
    (ncols, fcols), (nst, fst) = ctx.cloud(x, y, z).run(stages)
    tree = oracle.kd3(x, y, z)
    on, ons = oracle.stage_run(tree, ol.stage(ol.STAGE_NORMAL, 16), threads=8)
    of, ofs = oracle.stage_run(tree, ol.stage(ol.STAGE_COVFEAT, 16), threads=8)
    ok = ((nst & 0x3C) == 0) & (ons == 0)
    # normals agree with PDAL's up to the float-narrowing of its covariance
    dot = np.abs(ncols[0] * on[0] + ncols[1] * on[1] + ncols[2] * on[2])
    assert np.percentile(dot[ok], 1) > 1 - 1e-6
    assert np.all(np.sign(ncols[2][ok]) >= 0)
    e = rel_err(ncols[3][ok], on[3][ok], 1e-9)
    assert np.percentile(e, 99) < 1e-4


# ------------------------------------------------- full-size configurations --
def _full_check(ctx, oracle, x, y, z, stages, knn, sample):
    cl = ctx.cloud(x, y, z)
    ids, st = cl.knn(knn + 1)
    tree = oracle.kd3(x, y, z)
    ref = tree.neighbors_all(knn + 1, threads=16)
    eq = set_equal_rows(ids, ref)
    assert np.all(eq), f"{np.count_nonzero(~eq)} rows differ from KD3Index"
    flagged = (st & (g.ST_TIE | g.ST_INCOMPLETE)) != 0
    assert np.array_equal(ids[flagged], ref[flagged])
    res_cols, res_st = cl.run(stages, init=np.nan)
    rows = np.random.default_rng(0).choice(len(x), size=min(sample, len(x)), replace=False)
    rows.sort()
    for st_, cols, status in zip(stages, res_cols, res_st):
        ost = ol.stage(st_.stage, st_.knn, st_.stride, st_.mode, st_.features, st_.always_up,
                       st_.normalize)
        tc, ts, gap, ratio = oracle.truth_from_lists(x, y, z, ref[rows], ost, threads=16,
                                                     with_ratio=True)
        ok = ((status[rows] & 0x3C) == 0) & (ts == 0) & (gap > 1e-3)
        assert np.count_nonzero(ok) > 0.9 * len(rows)
        errs = truth_errors(ost, [c[rows] for c in cols], tc, ok, ratio)
        assert max(errs.values()) <= 1.0, errs


def test_config_plain_cpu_runnable_full(ctx, oracle):
    """250,000 points over 500 m x 500 m, covariancefeatures knn=10 with the
    ten non-density features: every row's neighbor set equals KD3Index's,
    features within the 1e-5 bar on a 50k-row sample."""
    x, y, z = g.synth_cloud(250_000, tile=0, side=500.0)
    _full_check(ctx, oracle, x, y, z, [g.stage(g.STAGE_COVFEAT, 10, features=g.F_ALL10)], 10,
                50_000)


def test_config_plain_gpu_full(ctx, oracle):
    """4,000,000 points, normal + covariancefeatures knn=16: every row's
    neighbor set equals KD3Index's; columns within the 1e-5 bar on a 100k-row
    sample."""
    x, y, z = g.synth_cloud(4_000_000, tile=0, side=500.0)
    _full_check(ctx, oracle, x, y, z,
                [g.stage(g.STAGE_NORMAL, 16), g.stage(g.STAGE_COVFEAT, 16)], 16, 100_000)


def test_config_precision_stress_full(ctx, oracle):
    """20,000,000 points (80 pts/m^2) with 5% exactly planar roof patches and
    2% exactly collinear wires, filters.eigenvalues + filters.covariancefeatures
    knn=32: neighbor sets equal KD3Index's on every row the search flagged and
    on a 2,000,000-row random sample (order too on flagged rows); every
    column within the 1e-5 bar of the extended-precision truth on a
    250,000-row sample -- relative on well-conditioned rows, absolute bounds
    on degenerate rows (smallest eigenvalue at rounding level)."""
    n, knn = 20_000_000, 32
    x, y, z = g.synth_cloud(n, tile=3, side=float(np.sqrt(n / 80.0)), roof_frac=0.05,
                            wire_frac=0.02)
    cl = ctx.cloud(x, y, z)
    ids, st = cl.knn(knn + 1)
    flagged = np.nonzero((st & (g.ST_TIE | g.ST_INCOMPLETE)) != 0)[0]
    assert len(flagged) > 0  # the roofs and wires do produce boundary ties
    rng = np.random.default_rng(33)
    sample = np.union1d(rng.choice(n, 2_000_000, replace=False), flagged)
    tree = oracle.kd3(x, y, z)
    ref = tree.neighbors_rows(sample, knn + 1, threads=16)
    eq = set_equal_rows(ids[sample], ref)
    assert np.all(eq), f"{np.count_nonzero(~eq)} sampled rows differ from KD3Index"
    fl = np.isin(sample, flagged)
    assert np.array_equal(ids[sample[fl]], ref[fl]), "re-solved rows keep KD3 order"
    del ids, ref
    stages = [g.stage(g.STAGE_EIGENVALUES, knn), g.stage(g.STAGE_COVFEAT, knn)]
    res_cols, res_st = cl.run(stages, init=np.nan)
    cl.close()
    rows = np.sort(rng.choice(n, 250_000, replace=False))
    nbr = tree.neighbors_rows(rows, knn + 1, threads=16)
    degenerate = 0
    for st_, cols, status in zip(stages, res_cols, res_st):
        ost = ol.stage(st_.stage, st_.knn, st_.stride, st_.mode, st_.features, st_.always_up,
                       st_.normalize)
        tc, ts, gap, ratio = oracle.truth_from_lists(x, y, z, nbr, ost, threads=16,
                                                     with_ratio=True)
        done = ((status[rows] & 0x3C) == 0) & (ts == 0)
        assert np.count_nonzero(done) > 0.95 * len(rows)
        # Well-conditioned rows need a resolved eigengap; degenerate rows
        # (flat roofs, wires) are checked absolutely whatever their gap.
        ok = done & ((gap > 1e-3) | (ratio <= 1e-10))
        degenerate = max(degenerate, int(np.count_nonzero(ok & (ratio <= 1e-10))))
        errs = truth_errors(ost, [c[rows] for c in cols], tc, ok, ratio)
        assert max(errs.values()) <= 1.0, errs
    assert degenerate > 0.03 * len(rows)  # roofs and wires are in the sample
