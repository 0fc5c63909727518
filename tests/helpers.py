"""Shared helpers for the parity tests."""
import numpy as np

import oracle_lib as ol


def small_cloud(n, seed=0, side=None, quant=0.01, dup_frac=0.0):
    """A quantized ALS-like cloud from the product generator, optionally with
    exact duplicates appended (duplicate coordinates create d2 == 0 ties)."""
    import gpupdal_amd as g
    side = side if side is not None else max(5.0, np.sqrt(n / 16.0))
    x, y, z = g.synth_cloud(n, tile=seed, side=side, quantum=quant)
    if dup_frac > 0:
        rng = np.random.default_rng(seed)
        m = int(n * dup_frac)
        idx = rng.integers(0, n, m)
        x = np.concatenate([x, x[idx]])
        y = np.concatenate([y, y[idx]])
        z = np.concatenate([z, z[idx]])
    return x, y, z


def set_equal_rows(a, b):
    """Row-wise set equality of two (n, k) id arrays."""
    return np.all(np.sort(a, axis=1) == np.sort(b, axis=1), axis=1)


def rel_err(a, b, floor):
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def boundary_tie_rows(oracle, x, y, z, kk, rows=None):
    """Rows whose kk-th and (kk+1)-th smallest squared distances are equal
    (brute force, so only for small clouds)."""
    ids, d2 = oracle.brute_knn(x, y, z, kk + 1, queries=rows)
    return d2[:, kk - 1] == d2[:, kk]


# Parity bar for stage columns against the extended-precision truth.
RTOL = 1e-5
# A row whose smallest eigenvalue is below this fraction of the largest has
# it at FP64 rounding level (exactly planar roofs, collinear wires): the
# columns that depend on it are then checked absolutely.
DEGENERATE_RATIO = 1e-10


def column_names(st):
    """Names of a stage's output columns, in order."""
    if st.stage == ol.STAGE_NORMAL:
        return ["NormalX", "NormalY", "NormalZ", "Curvature"]
    if st.stage == ol.STAGE_EIGENVALUES:
        return ["Eigenvalue0", "Eigenvalue1", "Eigenvalue2"]
    return [ol.FEATURE_NAMES[b] for b in range(10) if st.features & (1 << b)]


def truth_errors(st, cols, tcols, ok, ratio):
    """Per column: the largest error against the truth, as a multiple of the
    allowed error, over the rows in ok (1.0 = at the bar).

    Well-resolved rows (ratio > DEGENERATE_RATIO): |gpu - truth| <=
    RTOL * max(|truth|, 1e-6).  Degenerate rows: Omnivariance is skipped
    (the 1/6th power of a rounding-level eigenvalue is ~1e-3 in any FP64
    arithmetic); other columns must be within max(RTOL*|truth|, atol), with
    atol = 1e-7 for the unit-scale features (sqrt of a 1e-16 relative
    eigenvalue is 1.5e-8) and 1e-12 of the row's largest eigenvalue for the
    Eigenvalue columns.  In sqrt mode SurfaceVariation is sqrt(lambda_min)
    over the sum of the raw eigenvalues, so that noise scales as
    1 / sqrt(sum): its atol is 1e-7 / sqrt(EigenvalueSum) where the truth
    has that column (tiny neighborhoods, k = 3 on a 1 cm lattice, have sums
    near 1e-4 m^2).  Normals whose truth is horizontal are compared
    unsigned: always_up does not fix their sign."""
    out = {}
    names = column_names(st)
    deg = ok & (ratio <= DEGENERATE_RATIO)
    good = ok & ~deg
    lam_max = None
    if st.stage == ol.STAGE_EIGENVALUES:
        lam_max = np.abs(tcols[2])
    eigsum = None
    if st.stage == ol.STAGE_COVFEAT and "EigenvalueSum" in names:
        eigsum = np.abs(tcols[names.index("EigenvalueSum")])
    # A horizontal normal (nz = 0, a vertical neighborhood) has no "up":
    # always_up leaves its sign to the solver.
    flat_z = None
    if st.stage == ol.STAGE_NORMAL:
        flat_z = np.abs(tcols[2]) < 1e-9
    for name, c, t in zip(names, cols, tcols):
        if st.stage == ol.STAGE_NORMAL and name.startswith("Normal"):
            if not st.always_up:
                c, t = np.abs(c), np.abs(t)  # eigenvector sign is the solver's choice
            else:
                c = np.where(flat_z, np.abs(c), c)
                t = np.where(flat_z, np.abs(t), t)
        e = np.zeros(len(c))
        e[good] = np.abs(c[good] - t[good]) / (RTOL * np.maximum(np.abs(t[good]), 1e-6))
        if name != "Omnivariance":
            if lam_max is not None:
                atol = 1e-12 * lam_max[deg]
            elif name == "Curvature":
                atol = np.full(np.count_nonzero(deg), 1e-12)
            elif (name == "SurfaceVariation" and st.mode == ol.MODE_SQRT
                  and eigsum is not None):
                atol = 1e-7 / np.sqrt(np.minimum(np.maximum(eigsum[deg], 1e-300), 1.0))
            else:
                atol = np.full(np.count_nonzero(deg), 1e-7)
            allow = np.maximum(RTOL * np.abs(t[deg]), atol)
            e[deg] = np.abs(c[deg] - t[deg]) / np.maximum(allow, 1e-300)
        out[name] = float(np.nanmax(e)) if len(e) else 0.0
    return out
