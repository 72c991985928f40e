"""The C-ABI library loads and exports every symbol include/knng.h declares (no GPU needed)."""
import ctypes
import os
import re

import pytest

import sigmod23ann_amd as K

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "knng.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(knng_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(K.LIB_PATH):
        pytest.skip("libknng.so not built")
    L = ctypes.CDLL(K.LIB_PATH)
    declared = _declared()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(K.EXPORTS) == declared


def test_default_params_and_error_paths_without_gpu():
    if not os.path.exists(K.LIB_PATH):
        pytest.skip("libknng.so not built")
    p = K.Params()
    K.lib().knng_default_params(ctypes.byref(p))
    assert (p.k, p.leaf_max, p.small_max, p.iterations) == (100, 400, 3000, 1)
    # argument validation happens before any device call
    bad = K.params(k=0)
    h = ctypes.c_void_p()
    assert K.lib().knng_create(ctypes.byref(bad), ctypes.byref(h)) == K.ERR_INVALID
    assert b"k must be" in K.lib().knng_last_error()
    assert K.lib().knng_create(None, ctypes.byref(h)) == K.ERR_INVALID


def test_recall_grader_semantics():
    import numpy as np
    if not os.path.exists(K.LIB_PATH):
        pytest.skip("libknng.so not built")
    truth = np.array([[1, 2, 3, 4], [5, 6, 7, 8]], np.uint32)
    res = np.array([[4, 3, 9, 9], [8, 7, 6, 5]], np.uint32)   # 2/4 and 4/4
    assert K.recall(res, truth) == 0.75


def test_server_admission_and_validation_without_gpu():
    import numpy as np
    if not os.path.exists(K.LIB_PATH):
        pytest.skip("libknng.so not built")
    prm = K.params(k=100)
    f1 = K.job_footprint(100_000, 100, prm)
    f2 = K.job_footprint(1_000_000, 100, prm)
    f3 = K.job_footprint(1_000_000, 100, K.params(k=10))
    assert f2 > f1 > 0 and f3 < f2
    # per point: points (4 ld) and their fp16 copy for the side test (128 halves) + lists, outputs
    # (8k each), the snapshot and the reverse lists (8k; their sort works in the candidate buffers)
    # + two sets of NoN buffers, then per-point scalars and split tables (a few hundred bytes at most)
    core = 4 * 100 + 2 * 128 + 24 * 100 + 16 * 512
    assert 900_000 * core < f2 - f1 < 900_000 * (core + 200)
    assert f2 - K.job_footprint(1_000_000, 100, K.params(k=100, flags=K.FLAG_SIDE_FP32)) == 256 * 1_000_000
    # the KD splitter's projected rows (32 floats) in place of the fp16 rows, and the incremental
    # passes' second snapshot
    assert K.job_footprint(1_000_000, 100, K.params(k=100, flags=K.FLAG_KD_SPLIT)) - f2 == (128 - 256) * 1_000_000
    assert (K.job_footprint(1_000_000, 100, K.params(k=100, flags=K.FLAG_INCREMENTAL_NON)) - f2
            == 4 * 100 * 1_000_000)
    # a job larger than the budget is refused before anything is queued
    srv = K.Server(workers=2, hbm_budget=f1)
    X = np.zeros((1000, 100), np.float32)
    with pytest.raises(K.KnngError) as e:
        srv.submit(K.params(k=100), X=np.zeros((200_000, 100), np.float32))
    assert e.value.code == K.ERR_NOMEM
    with pytest.raises(K.KnngError) as e:
        srv.submit(K.params(k=0), X=X)
    assert e.value.code == K.ERR_INVALID
    srv.close()
    h = ctypes.c_void_p()
    assert K.lib().knng_server_create(0, 1 << 30, ctypes.byref(h)) == K.ERR_INVALID


def test_server_keeps_job_buffers_after_the_ticket_is_dropped():
    # The worker writes a job's ids into the submitted arrays whenever it runs: the server, not
    # the ticket, must keep them (and the host points) alive until the job is collected.
    import gc
    import weakref

    import numpy as np
    if not os.path.exists(K.LIB_PATH):
        pytest.skip("libknng.so not built")
    srv = K.Server(workers=1, hbm_budget=1 << 40)
    X = np.random.default_rng(0).normal(size=(500, 16)).astype(np.float32)
    t = srv.submit(K.params(k=10, leaf_max=64, rounds=1, non_passes=0), X=X)
    jid = t.id
    ids_ref = weakref.ref(srv._jobs[jid][0])
    x_ref = weakref.ref(srv._jobs[jid][2])
    del t, X
    gc.collect()
    assert ids_ref() is not None and x_ref() is not None
    srv.close()                       # drains the job (it fails without a GPU), then lets go
    assert srv._jobs == {}
    gc.collect()
    assert ids_ref() is None and x_ref() is None
