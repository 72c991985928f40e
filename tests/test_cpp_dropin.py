"""The C++ drop-in header compiles against a reference-style caller and, on the GPU, produces the
same graph as the Python path over the C ABI."""
import os
import subprocess

import numpy as np
import pytest

import sigmod23ann_amd as K

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _build(tmp_path):
    exe = tmp_path / "dropin"
    cmd = ["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "dropin_main.cpp"), "-o", str(exe),
           "-L", os.path.dirname(K.LIB_PATH), "-lknng",
           "-Wl,-rpath," + os.path.dirname(K.LIB_PATH)]
    subprocess.run(cmd, check=True)
    return exe


def test_dropin_header_compiles_and_links(tmp_path):
    if not os.path.exists(K.LIB_PATH):
        pytest.skip("libknng.so not built")
    assert _build(tmp_path).exists()


@pytest.mark.gpu
def test_dropin_runs_like_python(tmp_path):
    exe = _build(tmp_path)
    rng = np.random.default_rng(3)
    c = rng.uniform(-10, 10, size=(20, 100)).astype(np.float32)
    X = (c[rng.integers(0, 20, 3000)] + rng.normal(size=(3000, 100))).astype(np.float32)
    inp = tmp_path / "points.bin"
    with open(inp, "wb") as f:
        f.write(np.uint32(3000).tobytes())
        f.write(X.tobytes())
    out = tmp_path / "output.bin"
    subprocess.run([str(exe), str(inp), str(out)], check=True)
    got = np.fromfile(out, dtype=np.uint32).reshape(3000, 100)
    ref = K.Builder(K.params(k=100, leaf_max=256, rounds=4, non_passes=1, seed=7)).build(X)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.gpu
def test_dropin_random_kd_runs_like_python(tmp_path):
    # SolutionRandomKDMI355X: the reference's KD configuration (5 rounds, leaves < 400, no NoN)
    exe = _build(tmp_path)
    rng = np.random.default_rng(4)
    c = rng.uniform(-10, 10, size=(20, 100)).astype(np.float32)
    X = (c[rng.integers(0, 20, 3000)] + rng.normal(size=(3000, 100))).astype(np.float32)
    inp = tmp_path / "points.bin"
    with open(inp, "wb") as f:
        f.write(np.uint32(3000).tobytes())
        f.write(X.tobytes())
    out = tmp_path / "output.bin"
    subprocess.run([str(exe), str(inp), str(out), "kd"], check=True)
    got = np.fromfile(out, dtype=np.uint32).reshape(3000, 100)
    ref = K.Builder(K.params(k=100, leaf_max=399, rounds=5, non_passes=0, seed=7,
                             flags=K.FLAG_KD_SPLIT)).build(X)
    np.testing.assert_array_equal(got, ref)
