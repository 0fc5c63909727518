"""Writes the golden fixtures under tests/golden/ from the reference's own
nanoflann 1.3.1 and Eigen 3.4.0, compiled from the reference tree into
oracle/_ref/libref.so by `make -C oracle ref` (oracle/ref/ref_driver.cpp is
the harness).  The fixtures are data only: seeded inputs and the reference's
outputs on them.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as ol  # noqa: E402


def tie_cloud(rng, n):
    """ALS-like points on a 0.01 m lattice with many exact distance ties:
    a quantized plane patch, a few duplicated points and a vertical pole."""
    x = np.round(rng.uniform(0, 12, n) / 0.05) * 0.05 + 500000.0
    y = np.round(rng.uniform(0, 12, n) / 0.05) * 0.05 + 4000000.0
    z = np.round((0.02 * (x - 500000.0) + rng.normal(0, 0.03, n)) / 0.01) * 0.01
    d = n // 20
    x[:d], y[:d], z[:d] = x[d:2 * d], y[d:2 * d], z[d:2 * d]
    p = n // 40
    x[-p:], y[-p:] = 500006.0, 4000006.0
    z[-p:] = np.arange(p) * 0.25
    return x, y, z


def main():
    ref = ol.Ref()
    rng = np.random.default_rng(20240601)

    # KD3Index neighbor lists (query = each point's own coordinates).
    n = 2000
    x, y, z = tie_cloud(rng, n)
    tree = ref.kd3(x, y, z)
    lists = {}
    for k, stride in ((8, 1), (17, 1), (9, 2)):
        ids = np.zeros((n, k), np.uint16)
        for i in range(n):
            ids[i] = tree.neighbors(x[i], y[i], z[i], k, stride)
        lists[f"ids_k{k}_s{stride}"] = ids
    e = tree.export()
    np.savez_compressed(os.path.join(HERE, "kd3_ties.npz"), x=x, y=y, z=z,
                        vind=e["vind"].astype(np.uint32), leaf=e["leaf"],
                        divfeat=e["divfeat"], divlow=e["divlow"], divhigh=e["divhigh"],
                        **lists)

    # SelfAdjointEigenSolver<Matrix3d> on covariance-like matrices,
    # degenerate ones included.
    m = 600
    mats = np.zeros((m, 3, 3))
    for i in range(m):
        a = rng.normal(size=(3, 3))
        c = a @ a.T
        if i % 6 == 1:
            c[0, 2] = c[2, 0] = 0.0
        if i % 6 == 2:
            c = np.diag(rng.uniform(0, 1, 3))
        if i % 6 == 3:
            u = rng.normal(size=3)
            c = np.outer(u, u)
        if i % 6 == 4:
            u, v = rng.normal(size=3), rng.normal(size=3)
            c = np.outer(u, u) + np.outer(v, v)
        if i % 9 == 5:
            c *= 10.0 ** rng.integers(-9, 4)
        mats[i] = c
    ev = np.zeros((m, 3))
    vec = np.zeros((m, 3, 3))
    info = np.zeros(m, np.int32)
    for i in range(m):
        info[i], ev[i], vec[i] = ref.eigen_sym3(mats[i])
    np.savez_compressed(os.path.join(HERE, "eigen3.npz"), mats=mats, ev=ev, vec=vec,
                        info=info)

    # PDAL covariance (MathUtils.cpp:139-192) on neighbor lists of a
    # georeferenced cloud.
    n2 = 400
    cx = rng.uniform(0, 20, n2) + 500000.0
    cy = rng.uniform(0, 20, n2) + 4000000.0
    cz = rng.uniform(0, 5, n2) + 100.0
    sizes = rng.integers(3, 64, 300)
    ids = [rng.choice(n2, s, replace=False) for s in sizes]
    flat = np.concatenate(ids).astype(np.uint32)
    cov = np.array([ref.pdal_covariance(cx, cy, cz, i.astype(np.uint64)) for i in ids])
    np.savez_compressed(os.path.join(HERE, "covariance.npz"), x=cx, y=cy, z=cz,
                        sizes=sizes.astype(np.uint32), ids=flat, cov=cov)
    print("wrote kd3_ties.npz, eigen3.npz, covariance.npz")


if __name__ == "__main__":
    main()
