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
