"""The oracle against golden fixtures written by the reference's own
nanoflann 1.3.1 / Eigen 3.4.0 (tests/golden/make_golden.py): KD3Index
neighbor lists on a cloud full of exact distance ties, the tree itself,
SelfAdjointEigenSolver<Matrix3d> results and PDAL's covariance, all
bit for bit.  Runs on the CPU."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LIST_KEYS = (("ids_k8_s1", 8, 1), ("ids_k17_s1", 17, 1), ("ids_k9_s2", 9, 2))


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.fixture(scope="module")
def kd():
    return load("kd3_ties.npz")


def test_fixture_has_boundary_ties(oracle, kd):
    # The fixture is only worth its name if many rows have a distance tie
    # across the k-th neighbor boundary.
    x, y, z = kd["x"], kd["y"], kd["z"]
    _, d2 = oracle.brute_knn(x, y, z, 18)
    ties = np.sum(d2[:, 16] == d2[:, 17])
    assert ties > 50, ties


@pytest.mark.parametrize("key,k,stride", LIST_KEYS)
def test_kd3_lists_match_reference(oracle, kd, key, k, stride):
    x, y, z = kd["x"], kd["y"], kd["z"]
    tree = oracle.kd3(x, y, z)
    got = tree.neighbors_all(k, stride)
    want = kd[key].astype(np.uint32)
    bad = np.nonzero(np.any(got != want, axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} rows differ, first {bad[:5]}"


def test_kd3_tree_matches_reference(oracle, kd):
    e = oracle.kd3(kd["x"], kd["y"], kd["z"]).export()
    assert np.array_equal(e["vind"].astype(np.uint32), kd["vind"])
    leaf = (e["child1"] < 0).astype(np.int32)
    assert np.array_equal(leaf, kd["leaf"])
    inner = leaf == 0
    assert np.array_equal(e["divfeat"][inner], kd["divfeat"][inner])
    assert np.array_equal(e["divlow"][inner], kd["divlow"][inner])
    assert np.array_equal(e["divhigh"][inner], kd["divhigh"][inner])


def test_eigen_matches_reference_bits(oracle):
    g = load("eigen3.npz")
    for i in range(len(g["mats"])):
        info, ev, vec = oracle.eigen_sym3(g["mats"][i])
        assert info == g["info"][i]
        assert np.array_equal(ev, g["ev"][i]), i
        assert np.array_equal(vec, g["vec"][i]), i


def test_covariance_matches_reference_bits(oracle):
    g = load("covariance.npz")
    off = 0
    for i, s in enumerate(g["sizes"]):
        ids = g["ids"][off:off + s]
        off += s
        c = oracle.pdal_covariance(g["x"], g["y"], g["z"], ids)
        assert np.array_equal(c, g["cov"][i]), i
