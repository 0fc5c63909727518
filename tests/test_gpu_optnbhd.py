"""filters.optimalneighborhood on the GPU against the oracle: OptimalKNN and
OptimalRadius per row, bit for bit, on clouds with degenerate patches and on
a lattice whose lists are full of equal distances (KD3Index's order among
them decides every prefix).  The device's log may differ from the C
library's in the last bit, so a row whose two best entropies are within
1e-12 of each other (the oracle reports the gap) may choose the other k."""
import numpy as np
import pytest

import gpupdal_amd as g
from helpers import small_cloud

pytestmark = pytest.mark.gpu


def _check(ctx, oracle, x, y, z, kmin, kmax):
    kopt, ropt = ctx.cloud(x, y, z).optimal_neighborhood(kmin, kmax)
    okopt, oropt, gap, failed = oracle.optimal_neighborhood(oracle.kd3(x, y, z), kmin, kmax,
                                                            threads=8)
    assert failed == 0
    diff = (kopt != okopt) | (ropt.view(np.uint64) != oropt.view(np.uint64))
    assert np.all(gap[diff] < 1e-12), np.nonzero(diff)[0][:10]
    assert np.count_nonzero(diff) <= len(x) // 10000 + 1
    return kopt


@pytest.mark.parametrize("seed,n,dup", [(81, 20000, 0.05), (82, 60000, 0.0)])
def test_optimal_neighborhood_vs_oracle(ctx, oracle, seed, n, dup):
    x, y, z = small_cloud(n, seed=seed, dup_frac=dup)
    kopt = _check(ctx, oracle, x, y, z, 10, 14)
    assert np.count_nonzero((kopt >= 10) & (kopt <= 14)) > n // 2


def test_optimal_neighborhood_lattice_ties(ctx, oracle):
    g1 = np.arange(24, dtype=np.float64)
    X, Y, Z = np.meshgrid(g1, g1, np.arange(5, dtype=np.float64), indexing="ij")
    x = X.ravel() * 0.5 + 500000.0 + 0.013 * Z.ravel()
    y = Y.ravel() * 0.5 + 4000000.0
    z = Z.ravel() * 0.5 + 0.07 * X.ravel()
    _check(ctx, oracle, x, y, z, 8, 20)


def test_optimal_neighborhood_arguments(ctx):
    x, y, z = small_cloud(100, seed=3)
    cl = ctx.cloud(x, y, z)
    for a, b in ((1, 14), (15, 14), (10, 64)):
        with pytest.raises(g.GfError):
            cl.optimal_neighborhood(a, b)
    with pytest.raises(g.GfError):
        ctx.cloud(x[:9], y[:9], z[:9]).optimal_neighborhood(5, 10)
