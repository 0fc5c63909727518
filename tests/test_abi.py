"""The C ABI library without a GPU: it loads, exports every function
include/gpupdal_amd.h declares, and its host-only entries (stage column
counts, the seeded generator) behave.  No compute call needs a device here."""
import ctypes
import os
import re

import numpy as np
import pytest

import gpupdal_amd as g

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gpupdal_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_documented_entries():
    names = declared_functions()
    for must in ("gf_ctx_create", "gf_cloud_create", "gf_run", "gf_knn", "gf_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(g.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_list_matches_header():
    assert sorted(g.EXPORTED) == declared_functions()


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError, match="not built"):
        g.load_library(str(tmp_path / "libgpupdal_amd.so"))


def test_stage_columns():
    assert g.stage_columns(g.stage(g.STAGE_NORMAL, 8)) == 4
    assert g.stage_columns(g.stage(g.STAGE_EIGENVALUES, 8)) == 3
    assert g.stage_columns(g.stage(g.STAGE_COVFEAT, 10)) == 4
    assert g.stage_columns(g.stage(g.STAGE_COVFEAT, 10, features=g.F_ALL10)) == 10
    st = g.GfStage(g.STAGE_COVFEAT, 10, 1, g.MODE_SQRT, 1 << 12, 1, 0)
    assert g.lib().gf_stage_columns(ctypes.byref(st)) < 0  # unknown feature bit


def test_generator_is_deterministic_and_shaped():
    a = g.synth_cloud(20000, tile=3, side=50.0)
    b = g.synth_cloud(20000, tile=3, side=50.0)
    c = g.synth_cloud(20000, tile=4, side=50.0)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    assert not np.array_equal(a[0], c[0])
    x, y, z = a
    assert x.min() >= 500000.0 and x.max() <= 500050.0 + 1e-6
    assert y.min() >= 4000000.0 and y.max() <= 4000050.0 + 1e-6
    # LAS quantization: every coordinate is a multiple of 0.01.
    for v in (x - 500000.0, y - 4000000.0, z):
        q = v / 0.01
        assert np.max(np.abs(q - np.round(q))) < 1e-6


def test_tile_seed_is_splitmix_of_base_xor_tile():
    def splitmix64(v):
        m = (1 << 64) - 1
        v = (v + 0x9E3779B97F4A7C15) & m
        v = ((v ^ (v >> 30)) * 0xBF58476D1CE4E5B9) & m
        v = ((v ^ (v >> 27)) * 0x94D049BB133111EB) & m
        return v ^ (v >> 31)
    L = g.lib()
    for t in (0, 1, 255):
        assert L.gf_tile_seed(0x5EED, t) == splitmix64(0x5EED ^ t)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_context_without_gpu_reports_an_error():
    with pytest.raises(g.GfError):
        g.Context(0)
