"""Serving form through the C ABI: jobs of different (n, d, k) run concurrently on one GPU, each
result identical to the same job built alone; admission by device footprint."""
import time

import numpy as np
import pytest

import sigmod23ann_amd as K

pytestmark = pytest.mark.gpu


def lowrank(n, d, seed, latent=24):
    rng = np.random.default_rng(seed)
    W = rng.normal(size=(latent, d))
    z = rng.normal(size=(n, latent)) / np.sqrt(1.0 + np.arange(latent))
    X = z @ W + 0.05 * rng.normal(size=(n, d))
    return (X / np.linalg.norm(X, axis=1, keepdims=True)).astype(np.float32)


JOBS = [(3000, 64, 10), (20000, 100, 100), (50000, 128, 10), (8000, 100, 100), (30000, 64, 100)]


def job_params(k, seed):
    return K.params(k=k, leaf_max=256, small_max=3000, iterations=1, rounds=3, non_passes=1, seed=seed)


def test_serving_results_same_alone_and_in_mix():
    data = [(lowrank(n, d, 100 + i), job_params(k, 7 + i)) for i, (n, d, k) in enumerate(JOBS)]
    alone = [K.Builder(prm).build(X, return_dists=True) for X, prm in data]
    srv = K.Server(workers=3, hbm_budget=64 << 30)
    try:
        for rep in range(2):
            order = list(range(len(data))) if rep == 0 else list(reversed(range(len(data))))
            tickets = {i: srv.submit(data[i][1], X=data[i][0], return_dists=True) for i in order}
            for i in order:
                ids, ds, res = srv.wait(tickets[i])
                assert res["status"] == K.OK
                np.testing.assert_array_equal(ids, alone[i][0])
                np.testing.assert_array_equal(ds, alone[i][1])
    finally:
        srv.close()


def test_serving_device_points_and_budget_serialises():
    # equal footprints and a budget of exactly one: admitted jobs never overlap
    data = [(lowrank(20000, 100, 200 + i), job_params(100, 3 + i)) for i in range(3)]
    alone = [K.Builder(prm).build(X) for X, prm in data]
    budget = K.job_footprint(20000, 100, data[0][1])
    srv = K.Server(workers=3, hbm_budget=budget)
    dev = []
    try:
        tickets, t_sub = [], []
        for X, prm in data:
            dX = K.DeviceArray(X.nbytes)
            dX.upload(X)
            dev.append(dX)
            t_sub.append(time.perf_counter())
            tickets.append(srv.submit(prm, device_ptr=dX.ptr, n=X.shape[0], d=X.shape[1]))
        spans = []
        for i, t in enumerate(tickets):
            ids, _, res = srv.wait(t)
            np.testing.assert_array_equal(ids, alone[i])
            start = t_sub[i] + res["queue_ms"] / 1e3
            spans.append((start, start + res["run_ms"] / 1e3))
        spans.sort()
        for (s0, e0), (s1, e1) in zip(spans, spans[1:]):
            assert s1 >= e0 - 2e-3, "jobs overlapped although only one fits the budget"
    finally:
        srv.close()
        for dX in dev:
            dX.free()
