"""The PDAL-compatible C++ stages (gpupdal_amd/pdal/libgfpdal.so) through
their C++ test program tests/cpp/stage_test: option parsing, dimension
registration and preconditions on the CPU; the upstream known-answer
clouds, skips and chaining on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "stage_test")


def run(suite):
    if not os.path.exists(BIN):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "gpupdal_amd", "pdal")], check=True)
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], check=True)
    p = subprocess.run([BIN, suite], capture_output=True, text=True, timeout=300)
    print(p.stdout)
    fails = [l for l in p.stdout.splitlines() if l.startswith("FAIL")]
    assert p.returncode == 0 and not fails, "\n".join(fails) or p.stderr


def test_stage_options_and_dims():
    run("Cpu")


@pytest.mark.gpu
def test_stages_on_gpu():
    run("Gpu")
