"""Contest file formats and the grader tool (no GPU): the points reader, the id-file writer, and
knng_grade against a restatement of src/grade.cpp:61-95 (baseline ids counted with repeats)."""
import os
import subprocess

import numpy as np
import pytest

import sigmod23ann_amd as K


def test_points_file_roundtrip(tmp_path):
    X = np.random.default_rng(1).normal(size=(37, 100)).astype(np.float32)
    p = tmp_path / "pts.bin"
    K.write_points(p, X)
    assert os.path.getsize(p) == 4 + 37 * 400
    np.testing.assert_array_equal(K.read_points(p), X)
    with open(p, "r+b") as f:          # header larger than the rows present
        f.write(np.uint32(38).tobytes())
    with pytest.raises(ValueError):
        K.read_points(p)


def grade_reference(base, test, n):
    total = 0
    for i in range(n):
        b = base[i] if i < len(base) else []
        t = set(test[i].tolist()) if i < len(test) else set()
        total += sum(1 for x in b if int(x) in t)
    return total / n


def test_grade_tool_matches_grade_semantics(tmp_path):
    if not os.path.exists(K.GRADE_PATH):
        pytest.skip("knng_grade not built")
    rng = np.random.default_rng(2)
    n = 50
    base = rng.integers(0, 400, size=(n, 100)).astype(np.uint32)
    base[3, :10] = 1                    # repeated ids (padding) count once per occurrence
    test = base.copy()
    test[:, ::3] = rng.integers(400, 800, size=test[:, ::3].shape)
    test_short = test[:45]              # missing rows score 0
    K.save_graph(tmp_path / "base.bin", base)
    K.save_graph(tmp_path / "test.bin", test_short)
    out = subprocess.run([K.GRADE_PATH, str(tmp_path / "base.bin"), str(tmp_path / "test.bin"), str(n)],
                         check=True, capture_output=True, text=True).stdout
    lines = out.strip().splitlines()
    assert lines[2] == f"num points: {n}"
    want = grade_reference(base, test_short, n)
    assert lines[3] == f"avg percent: {want:g}"


def test_cli_rejects_short_file_without_gpu(tmp_path):
    if not os.path.exists(K.CLI_PATH):
        pytest.skip("knng_cli not built")
    p = tmp_path / "pts.bin"
    K.write_points(p, np.zeros((10, 100), np.float32))
    with open(p, "r+b") as f:
        f.write(np.uint32(11).tobytes())
    r = subprocess.run([K.CLI_PATH, str(p), "--out", str(tmp_path / "o.bin")], capture_output=True, text=True)
    assert r.returncode == 1 and "file holds 10" in r.stderr
