"""Trees deeper than the exact re-solve's 64 register frames.  Points along a
line spaced geometrically (x_i = 1e-9 * 1.02^i, each point twice) make
nanoflann's middle split peel about 70 points per level, so 6000 points
build a KD3 tree of more than 80 levels.  The oracle's tree is that deep, and is the reference's
own nanoflann tree node for node (oracle/_ref), with the same neighbor lists;
the GPU side of this cloud is tests/test_gpu_deep_tree.py."""
import numpy as np
import pytest

import oracle_lib as ol


def skewed_line(n=3000, r=1.02):
    """n points at 1e-9 * r^i on the x axis, each twice: with an odd number
    of neighbors every row's boundary distance is tied, so every row goes
    to the exact search."""
    x = 1e-9 * r ** np.arange(n)
    x = np.concatenate([x, x])
    return x, np.zeros(2 * n), np.zeros(2 * n)


def tree_depth(c1, c2):
    best, st = 0, [(0, 1)]
    while st:
        nd, d = st.pop()
        best = max(best, d)
        if c1[nd] >= 0:
            st.append((c1[nd], d + 1))
        if c2[nd] >= 0:
            st.append((c2[nd], d + 1))
    return best


@pytest.fixture(scope="module")
def oracle():
    return ol.Oracle()


def test_skewed_line_tree_is_deeper_than_64(oracle):
    e = oracle.kd3(*skewed_line()).export()
    assert tree_depth(e["child1"], e["child2"]) > 80


@pytest.mark.skipif(not ol.ref_available(), reason="oracle/_ref/libref.so not built")
def test_skewed_line_tree_equals_reference(oracle):
    x, y, z = skewed_line()
    to, tr = oracle.kd3(x, y, z), ol.Ref().kd3(x, y, z)
    eo, er = to.export(), tr.export()
    assert np.array_equal(eo["vind"], er["vind"])
    leaf = eo["child1"] < 0
    assert np.array_equal(leaf.astype(np.int32), er["leaf"])
    for key in ("divfeat", "divlow", "divhigh"):
        assert np.array_equal(eo[key][~leaf], er[key][~leaf])
    for i in range(0, len(x), 37):
        for k in (15, 16):
            assert np.array_equal(to.neighbors(x[i], y[i], z[i], k),
                                  tr.neighbors(x[i], y[i], z[i], k))
