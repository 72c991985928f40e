"""Serving form through the C ABI: jobs of different (n, d, k) run concurrently on one GPU, each
result identical to the same job built alone; admission by device footprint, counting what idle
workers keep; a short Poisson stream of the serving workload with per-job recall."""
import threading
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


def _device_used():
    free, total = K.mem_get_info()
    return total - free


def test_serving_jobs_overlap_when_the_budget_allows():
    # Two jobs submitted together with room for both: two workers run them at once, and the
    # pair takes less wall time than the two builds one after the other.
    data = [(lowrank(150_000, 100, 300 + i), job_params(100, 11 + i)) for i in range(2)]
    b = K.Builder(data[0][1])
    b.build(data[0][0])                                   # load the kernels
    alone = []
    for X, prm in data:
        b.set_params(prm)
        t0 = time.perf_counter()
        alone.append((b.build(X), time.perf_counter() - t0))
    srv = K.Server(workers=2, hbm_budget=64 << 30)
    try:
        warm = [srv.submit(prm, X=X) for X, prm in data]  # both workers grow their buffers
        for t in warm:
            srv.wait(t)
        t0 = time.perf_counter()
        tickets = [srv.submit(prm, X=X) for X, prm in data]
        out = [srv.wait(t) for t in tickets]
        wall = time.perf_counter() - t0
    finally:
        srv.close()
    for (ids, _, res), (ref, _) in zip(out, alone):
        np.testing.assert_array_equal(ids, ref)
    (s0, e0), (s1, e1) = [(r["start_ms"], r["done_ms"]) for _, _, r in out]
    assert {r["worker"] for _, _, r in out} == {0, 1}
    assert max(s0, s1) < min(e0, e1), "admitted jobs did not run at the same time"
    serial = sum(t for _, t in alone)
    print(f"pair wall {wall * 1e3:.1f} ms, one after the other {serial * 1e3:.1f} ms")
    assert wall < serial


def test_serving_budget_of_one_large_job_with_alternating_sizes():
    # Budget = one large job's footprint, two workers, large and small jobs alternating: every job
    # succeeds with its alone result, the accounted memory never exceeds the budget, and neither
    # does the device memory actually in use (sampled while the jobs run).
    big = (lowrank(120_000, 100, 400), job_params(100, 21))
    small = (lowrank(15_000, 64, 401), job_params(10, 22))
    seq = [big, small, big, small, big, small]
    alone = {id(big): K.Builder(big[1]).build(big[0]), id(small): K.Builder(small[1]).build(small[0])}
    budget = K.job_footprint(120_000, 100, big[1])
    K.lib().knng_synchronize()
    base = _device_used()
    peak = [0]
    stop = threading.Event()

    def sample():
        while not stop.is_set():
            peak[0] = max(peak[0], _device_used() - base)
            time.sleep(0.002)

    th = threading.Thread(target=sample)
    srv = K.Server(workers=2, hbm_budget=budget)
    th.start()
    try:
        tickets = [(job, srv.submit(job[1], X=job[0])) for job in seq]
        for job, t in tickets:
            ids, _, res = srv.wait(t)
            assert res["status"] == K.OK
            assert res["worker_device_bytes"] <= budget
            np.testing.assert_array_equal(ids, alone[id(job)])
        st = srv.stats()
    finally:
        stop.set()
        th.join()
        srv.close()
    print(f"budget {budget / 2**20:.0f} MiB, accounted peak {st['peak_in_use'] / 2**20:.0f} MiB, "
          f"device peak {peak[0] / 2**20:.0f} MiB, trims {st['trims']}")
    assert st["peak_in_use"] <= budget
    assert st["trims"] >= 1                 # a small job followed a large one on a full budget
    assert peak[0] <= budget + (256 << 20)  # streams, events and allocation granularity


@pytest.mark.parametrize("flags", [0, K.FLAG_KD_SPLIT, K.FLAG_INCREMENTAL_NON])
def test_job_footprint_bounds_what_a_build_keeps(flags):
    # a fresh worker per job: what it keeps afterwards is what that one build allocated
    for n, d, k in [(40_000, 100, 100), (25_000, 128, 10), (60_000, 64, 100)]:
        prm = K.params(k=k, leaf_max=256, rounds=2, non_passes=2, seed=5, flags=flags,
                       search_k=max(k, 100) if k < 100 else 0)
        srv = K.Server(workers=1, hbm_budget=64 << 30)
        try:
            _, _, res = srv.wait(srv.submit(prm, X=lowrank(n, d, n + d), return_dists=True))
        finally:
            srv.close()
        assert 0 < res["worker_device_bytes"] <= K.job_footprint(n, d, prm), (n, d, k)


def _recall_device(X, got, k, queries):
    """recall@k of the rows `got` (for `queries`) against the FP64 exact k-NN kernel."""
    n, d = X.shape
    chk = K.Builder(K.params(k=k))
    dX = K.DeviceArray(X.nbytes)
    dX.upload(X)
    nq = queries.size
    dq = K.DeviceArray(nq * 4)
    dq.upload(queries)
    dl = K.DeviceArray(nq * k * 4)
    dl.upload(np.ascontiguousarray(got, np.uint32))
    dub = K.DeviceArray(nq * 8)
    chk.exact_upper_from_lists(dX.ptr, n, d, d, dq.ptr, nq, dl.ptr, k, dub.ptr)
    dti = K.DeviceArray(nq * k * 4)
    dtd = K.DeviceArray(nq * k * 8)
    chk.exact_knn_device(dX.ptr, n, d, d, dq.ptr, nq, k, dub.ptr, 4 * k + 256, dti.ptr, dtd.ptr)
    truth = dti.download((nq, k), np.uint32)
    for a in (dX, dq, dl, dub, dti, dtd):
        a.free()
    return K.recall(got, truth)


def test_poisson_stream_of_serving_jobs_each_above_recall():
    # The serving workload in small: seeded Poisson arrivals (10 jobs/s) of 20 jobs, N
    # log-uniform in [1e4, 2e5], d in {64, 100, 128}, k in {10, 100}, low-rank points with a
    # per-job seed; 4 workers; each job's recall@k on 2,000 sampled points >= 0.94.
    rng = np.random.default_rng(2024)
    srv = K.Server(workers=4, hbm_budget=250 << 30)
    jobs = []
    try:
        t_next = time.perf_counter()
        for i in range(20):
            r = np.random.default_rng([2024, i])
            n = int(round(np.exp(r.uniform(np.log(1e4), np.log(2e5)))))
            d, k = int(r.choice([64, 100, 128])), int(r.choice([10, 100]))
            X = lowrank(n, d, 5000 + i)
            prm = K.params(k=k, leaf_max=400, small_max=3000, iterations=1, rounds=6,
                           non_passes=2, seed=5000 + i, search_k=max(k, 100))
            t_next += rng.exponential(0.1)
            time.sleep(max(0.0, t_next - time.perf_counter()))
            jobs.append((X, k, srv.submit(prm, X=X)))
        results = [(X, k, srv.wait(t)) for X, k, t in jobs]
    finally:
        srv.close()
    lat = []
    for i, (X, k, (ids, _, res)) in enumerate(results):
        assert res["status"] == K.OK
        n = X.shape[0]
        q = np.sort(np.random.default_rng([i, 7]).choice(n, size=min(n, 2000), replace=False)).astype(np.uint32)
        rec = _recall_device(X, ids[q], k, q)
        lat.append(res["latency_ms"])
        print(f"job {i}: n={n} d={X.shape[1]} k={k} latency {res['latency_ms']:.0f} ms recall {rec:.4f}")
        assert rec >= 0.94, (i, n, k, rec)
    print(f"latency p50 {np.percentile(lat, 50):.0f} ms p99 {np.percentile(lat, 99):.0f} ms")
