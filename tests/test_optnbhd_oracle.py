"""The oracle's filters.optimalneighborhood (filters/OptimalNeighborhoodFilter.cpp:
72-163) pinned without the reference: the running mean / scatter recurrence
against a two-pass sample covariance of each neighbor prefix, the argmin
against numpy's eigvalsh on those covariances, and the degenerate cases the
formula fixes (collinear and coplanar neighborhoods have a zero eigenvalue,
so 0 * log 0 makes every entropy NaN and PDAL keeps OptimalKNN 0)."""
import numpy as np

from helpers import small_cloud


def _entropy(cov):
    ev = np.linalg.eigvalsh(cov)
    lam = np.maximum(ev[::-1], 0.0)
    lam = lam / lam.sum()
    with np.errstate(divide="ignore", invalid="ignore"):
        return -(lam[2] * np.log(lam[2]) + lam[1] * np.log(lam[1]) + lam[0] * np.log(lam[0]))


def test_oracle_against_two_pass_prefixes(oracle):
    """Where every prefix's smallest eigenvalue is well away from zero (so
    no entropy is NaN by rounding) and the best entropy is clearly the
    smallest, the oracle's choice is numpy's."""
    x, y, z = small_cloud(3000, seed=71)
    tree = oracle.kd3(x, y, z)
    kopt, ropt, gap, failed = oracle.optimal_neighborhood(tree, 10, 14, threads=4)
    assert failed == 0
    checked = 0
    for i in range(len(x)):
        ids = tree.neighbors(x[i], y[i], z[i], 14).astype(np.int64)
        pts = np.stack([x[ids], y[ids], z[ids]], axis=1)
        covs = [np.cov(pts[:k].T, ddof=1) for k in range(10, 15)]
        evs = [np.linalg.eigvalsh(c) for c in covs]
        if min(e[0] / e[2] for e in evs) < 1e-6:
            continue
        ents = np.array([_entropy(c) for c in covs])
        srt = np.sort(ents)
        if (srt[1] - srt[0]) / abs(srt[0]) < 1e-6:
            continue
        best = int(np.argmin(ents)) + 10
        assert kopt[i] == best, i
        d2 = ((x[ids] - x[i]) ** 2 + (y[ids] - y[i]) ** 2) + (z[ids] - z[i]) ** 2
        assert ropt[i] == np.sqrt(d2[best - 1])
        checked += 1
    assert checked > 1000


def test_oracle_degenerate_neighborhoods(oracle):
    t = np.arange(40, dtype=np.float64)
    line = (t * 0.5 + 100.0, t * 0.25 + 200.0, t * 0.0 + 5.0)
    tree = oracle.kd3(*line)
    kopt, ropt, _, failed = oracle.optimal_neighborhood(tree, 10, 14)
    assert failed == 0 and np.all(kopt == 0) and np.all(ropt == 0.0)
    g1 = np.arange(12, dtype=np.float64)
    X, Y = np.meshgrid(g1, g1)
    plane = (X.ravel() + 0.1 * Y.ravel(), Y.ravel(), np.full(X.size, 7.0))
    kopt, ropt, _, _ = oracle.optimal_neighborhood(oracle.kd3(*plane), 10, 14)
    assert np.all(kopt == 0) and np.all(ropt == 0.0)
