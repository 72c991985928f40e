"""knng_cli and knng_grade end to end on the GPU: the contest-format build equals the library
build, --truth equals the FP64 oracle, and the grader agrees with the library's recall."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import sigmod23ann_amd as K

pytestmark = pytest.mark.gpu


def lowrank(n, d, seed, latent=24):
    rng = np.random.default_rng(seed)
    W = rng.normal(size=(latent, d))
    z = rng.normal(size=(n, latent)) / np.sqrt(1.0 + np.arange(latent))
    X = z @ W + 0.05 * rng.normal(size=(n, d))
    return (X / np.linalg.norm(X, axis=1, keepdims=True)).astype(np.float32)


def test_cli_build_truth_and_grade(tmp_path):
    n = 6000
    X = lowrank(n, 100, 51)
    src = tmp_path / "points.bin"
    K.write_points(src, X)
    args = ["--rounds", "3", "--non", "1", "--leaf", "256", "--seed", "77"]
    subprocess.run([K.CLI_PATH, str(src), "--out", str(tmp_path / "output.bin")] + args,
                   check=True, capture_output=True, text=True)
    got = K.read_graph(tmp_path / "output.bin", n)
    want = K.Builder(K.params(k=100, rounds=3, non_passes=1, leaf_max=256, seed=77)).build(X)
    np.testing.assert_array_equal(got, want)

    subprocess.run([K.CLI_PATH, str(src), "--out", str(tmp_path / "truth.bin"), "--truth"] + args,
                   check=True, capture_output=True, text=True)
    truth = K.read_graph(tmp_path / "truth.bin", n)
    o_ids, o_d = oracle.exact_knn(X, np.arange(n, dtype=np.uint32), 100)
    assert (truth == o_ids).mean() > 0.9999      # only exact FP64 ties may order differently
    out = subprocess.run([K.GRADE_PATH, str(tmp_path / "truth.bin"), str(tmp_path / "output.bin"), str(n)],
                         check=True, capture_output=True, text=True).stdout
    avg = float(out.strip().splitlines()[-1].split(":")[1])
    assert abs(avg - 100 * K.recall(got, truth)) < 1e-3


def test_large_device_to_host_copy_through_pinned_stages():
    # knng_memcpy_d2h sends copies of 64 MB and more through two pinned 32 MB buffers: a size that
    # is neither a multiple of the buffer nor of the host copy's 4096-byte pieces, written on the
    # device by a copy of a host array, must come back unchanged, and so must a small copy
    n = (96 << 20) // 4 + 1237
    a = np.random.default_rng(3).integers(0, 2**32, size=n, dtype=np.uint32)
    d = K.DeviceArray(a.nbytes)
    d.upload(a)
    np.testing.assert_array_equal(d.download(a.shape, np.uint32), a)
    np.testing.assert_array_equal(d.download((1000,), np.uint32), a[:1000])
