"""Benchmark: KNN-graph build throughput on one MI355X.

Workload (default): 10,000,000 x 100 float32 points, low-rank embedding-like data (24 latent dims,
sigma_i = 1/sqrt(1+i), noise 0.05, unit norm), k = 100, leaf <= 400, 6 split rounds, 2
neighbours-of-neighbours passes, generated on the device so the
input is resident in HBM before timing. One step = one whole KNN-graph build (all split rounds,
leaf stages, the neighbours-of-neighbours pass and the finalize) through the C ABI
(knng_build_device). value = points / second of the build.

Also reported: recall@k against exact FP64 k-NN of sampled points, the roofline of the dominant
kernel (HIP-event timings from the library), and a CPU baseline: the CPU restatement of the
reference path (oracle/) timed on this host's cores on a bounded prefix of the same data.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "KNN-graph points/sec at recall@100≥0.94 (10M×100, k=100); p99 job latency"
HBM_PEAK_GBS = 8000.0
MFMA_FP16_PEAK_TFLOPS = 2500.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="lowrank", choices=["lowrank", "zipf128", "serving"],
                    help="lowrank: the headline 10M x 100 build; zipf128: one round (split + leaf "
                         "stage) over 10M x 128 Zipf-mixture points with buckets up to --leaf; "
                         "serving: a Poisson stream of build jobs through the job server")
    ap.add_argument("--rate", type=float, default=2.0, help="serving: jobs per second")
    ap.add_argument("--duration", type=float, default=60.0, help="serving: seconds of arrivals")
    ap.add_argument("--workers", type=int, default=4, help="serving: concurrent builders")
    ap.add_argument("--budget-gb", type=float, default=250.0, help="serving: HBM admission budget")
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--d", type=int, default=100)
    ap.add_argument("--k", type=int, default=100)
    # the smallest round count whose recall clears 0.94 with margin at 10M (2 NoN passes):
    # 5 rounds give 0.9432 / 0.9403 on two seeds, 6 give 0.9537 / 0.9512
    ap.add_argument("--rounds", type=int, default=6)
    ap.add_argument("--leaf", type=int, default=400)
    ap.add_argument("--iterations", type=int, default=1)
    ap.add_argument("--non", type=int, default=2)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp16", "bf16"])
    ap.add_argument("--splitter", default="kmeans", choices=["kmeans", "kd"],
                    help="kmeans: SolutionKmeans' 2-means split; kd: SolutionRandomKD's "
                         "random-rotation split (the reference runs it with 5 rounds, leaves < 400, "
                         "no NoN: --splitter kd --rounds 5 --leaf 399 --non 0)")
    ap.add_argument("--leaf-mode", default="binned", choices=["binned", "batched", "graph"],
                    help="leaf stage: one launch per bucket-size bin; one persistent kernel over the "
                         "round's work list; the same replayed from a captured hipGraph")
    ap.add_argument("--non-fp32-gram", action="store_true",
                    help="neighbours-of-neighbours screen on the FP32 Gram of the rows instead of the "
                         "fp16 Gram of the rows (each scaled by its power of two, exact margin)")
    ap.add_argument("--non-screen16", action="store_true",
                    help="neighbours-of-neighbours screen on a 16-bit Gram of rows centred on each "
                         "q instead of the FP32 Gram of the rows")
    ap.add_argument("--no-bucket-limit", action="store_true",
                    help="exact form: leaf screen against each point's list bound only, without the "
                         "in-bucket limit (the k-th smallest upper bound of its bucket distances)")
    ap.add_argument("--side-fp32", action="store_true",
                    help="split side test on the FP32 rows only (default: fp16 copy of the rows with "
                         "an exact error bound, the rows inside it decided on their FP32 values)")
    ap.add_argument("--autotune", action="store_true",
                    help="leaf workgroup shape picked by timing both shapes on the first launch of "
                         "each size bin (the warmup builds pay for it)")
    ap.add_argument("--seed", type=int, default=2023)
    ap.add_argument("--recall-queries", type=int, default=100_000)
    ap.add_argument("--cpu-sample", type=int, default=250_000)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--compare-fp32", action="store_true",
                    help="also build the plain FP32 GPU form (FP32 Gram in both stages) and, with "
                         "--precision bf16/fp16, the FP32 headline form, at the same rounds, and "
                         "report the recall deltas and the speed ratios")
    ap.add_argument("--error-queries", type=int, default=10_000,
                    help="sampled points whose returned distances are checked against FP64")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="before the build, run one plain device copy of this many MiB (a known "
                         "read/write volume for scaling FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_non_reverse_traffic_lowrank10M.md")


def traffic_key(a):
    """The workload a traffic figure belongs to: the build's sizes and its screen forms."""
    flags = [f for f, on in (("non-fp32-gram", a.non_fp32_gram), ("non-screen16", a.non_screen16))
             if on]
    return (f"{a.workload} n={a.n} d={a.d} k={a.k} rounds={a.rounds} leaf={a.leaf} non={a.non} "
            f"precision={a.precision} {a.splitter}" + "".join(" " + f for f in flags))


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed counter summary, if it covers this
    kernel on this workload (the summary names both; another size or screen form moves the
    figure, so it is then not reported)."""
    try:
        with open(TRAFFIC_FILE) as f:
            vals = dict(line.strip()[2:].split(": ", 1) for line in f
                        if line.startswith("- ") and ": " in line)
    except OSError:
        return None
    if vals.get("kernel") != kernel or vals.get("workload") != workload:
        return None
    try:
        return float(vals["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def copy_calibration(mib):
    """One elementwise copy of `mib` MiB on the device: mib MiB read and mib MiB written."""
    import torch
    x = torch.ones(mib * (1 << 18), dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    torch.add(x, 0.0, out=y)
    torch.cuda.synchronize()
    log(f"[bench] copy calibration: {mib} MiB read, {mib} MiB written")


def leaf_flags(a, K):
    return ({"binned": 0, "batched": K.FLAG_LEAF_BATCHED, "graph": K.FLAG_LEAF_GRAPH}[a.leaf_mode]
            | (K.FLAG_NO_BUCKET_LIMIT if a.no_bucket_limit else 0)
            | (K.FLAG_SIDE_FP32 if a.side_fp32 else 0))


def side_row_bytes(a, d):
    """Algorithmic bytes per side-test position: the row (FP32, or its fp16 copy by default), the
    permutation entry and the flag byte. The fp16 form's re-reads of FP32 rows inside its error
    bound are not counted."""
    return (4 * d if a.side_fp32 or a.splitter == "kd" else 2 * d) + 4 + 1


def zipf_round(a, K):
    """One round of the batching form: 2-means split + every bucket's all-pairs stage."""
    n, d, k = a.n, (a.d if a.d != 100 else 128), a.k
    leaf = a.leaf if a.leaf != 400 else 4096
    ld = (d + 3) & ~3
    prm = K.params(k=k, leaf_max=leaf, small_max=3000, iterations=a.iterations, rounds=1,
                   non_passes=0, precision=K.PREC_FP32, seed=a.seed,
                   flags=(K.FLAG_AUTOTUNE if a.autotune else 0) | leaf_flags(a, K))
    b = K.Builder(prm)
    log(f"[bench] generating {n}x{d} Zipf-mixture points (20,000 clusters) on the device")
    dX = K.DeviceArray(n * ld * 4)
    b.generate_device(K.GEN_ZIPF, n, d, ld, a.seed, dX.ptr)
    dIds = K.DeviceArray(n * k * 4)
    dD = K.DeviceArray(n * k * 4)
    times, stats = [], []
    for s in range(a.warmup + a.steps):
        K.lib().knng_synchronize()
        t0 = time.perf_counter()
        b.build_device(dX.ptr, n, d, ld, dIds.ptr, dD.ptr)
        K.lib().knng_synchronize()
        dt = time.perf_counter() - t0
        st = b.stats()
        log(f"[bench] zipf round {s}: {dt:.3f} s  split {st['split_ms']:.1f} ms (side {st['side_ms']:.1f}) "
            f"leaf {st['leaf_ms']:.1f} ms, {st['leaves']} buckets")
        if s >= a.warmup:
            times.append(dt)
            stats.append(st)
    st = stats[-1]
    mean_t = float(np.mean(times))
    side_bytes = st["side_rows"] * side_row_bytes(a, d)
    gram_flops = 2.0 * d * st["leaf_pairs"]
    out = {
        "metric": "one round (split + all buckets) of the batching form, 10M x 128 Zipf mixture",
        "value": round(mean_t * 1e3, 2), "unit": "ms per round", "higher_is_better": False,
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"Zipf(1.1) mixture {n}x{d}, 20000 clusters, sigma U[0.3,2], "
                               f"leaf<={leaf}, k={k}", "n": n, "d": d, "k": k, "leaf_max": leaf,
                   "leaf_mode": a.leaf_mode, "bucket_limit": not a.no_bucket_limit,
                   "side_test": "fp32" if a.side_fp32 else "fp16 with exact bound"},
        "stages_ms": {"split": round(st["split_ms"], 2), "side": round(st["side_ms"], 2),
                      "leaf": round(st["leaf_ms"], 2), "total": round(st["total_ms"], 2)},
        "buckets": int(st["leaves"]), "leaf_pairs": int(st["leaf_pairs"]),
        "split_side_hbm": {"achieved": round(side_bytes / (st["side_ms"] * 1e-3) / 1e9, 1),
                           "peak": HBM_PEAK_GBS, "unit": "GB/s"},
        "leaf_gram_mfma": {"achieved": round(gram_flops / (st["leaf_ms"] * 1e-3) / 1e12, 2),
                           "peak": MFMA_FP16_PEAK_TFLOPS, "unit": "TFLOP/s",
                           "frac": round(gram_flops / (st["leaf_ms"] * 1e-3) / 1e12 / MFMA_FP16_PEAK_TFLOPS, 5)},
    }
    print(json.dumps(out), flush=True)


def serving(a, K):
    """Continuous batching: Poisson arrivals of build jobs (N log-uniform in [1e4, 2e6], d in
    {64, 100, 128}, k in {10, 100}, low-rank points from a per-job seed, generated on the device
    at arrival) through knng_server_*; then recall@k of every job on a 10k-point FP64 sample."""
    import threading
    rng = np.random.default_rng(a.seed)
    n_max, d_max = 2_000_000, 128
    srv = K.Server(workers=a.workers, hbm_budget=int(a.budget_gb * (1 << 30)))
    gen = K.Builder(K.params(k=10))
    pool = [K.DeviceArray(n_max * d_max * 4) for _ in range(a.workers + 4)]
    free = list(range(len(pool)))
    cv = threading.Condition()
    results, threads = [], []

    def job_spec(i):
        r = np.random.default_rng([a.seed, i])
        n = int(round(np.exp(r.uniform(np.log(1e4), np.log(2e6)))))
        return n, int(r.choice([64, 100, 128])), int(r.choice([10, 100])), 1000 + i

    def finish(i, t, slot, n, d, k, seed, t_arr):
        ids, _, res = srv.wait(t)
        t_done = time.perf_counter()
        log(f"[bench] job {i} done: n={n} d={d} k={k} latency {(t_done - t_arr) * 1e3:.0f} ms")
        ns = min(n, 10_000)
        q = np.sort(np.random.default_rng([seed, 7]).choice(n, size=ns, replace=False)).astype(np.uint32)
        with cv:
            results.append(dict(i=i, n=n, d=d, k=k, seed=seed, q=q, got=np.ascontiguousarray(ids[q]),
                                arrival=t_arr, done=t_done, **res))
            free.append(slot)
            cv.notify_all()

    # warm each worker's buffers with the largest shape, so the stream does not pay for growth
    warm = []
    for w in range(a.workers):
        gen.generate_device(K.GEN_LOWRANK, n_max, d_max, d_max, 1, pool[w].ptr)
        warm.append(srv.submit(K.params(k=100, leaf_max=a.leaf, rounds=1, non_passes=1, seed=1),
                               device_ptr=pool[w].ptr, n=n_max, d=d_max))
    for t in warm:
        srv.wait(t)
    log(f"[bench] serving: {a.workers} workers warmed; Poisson arrivals at {a.rate}/s for {a.duration} s")
    t0 = time.perf_counter()
    t_next, i = t0 + rng.exponential(1.0 / a.rate), 0
    while t_next - t0 < a.duration:
        now = time.perf_counter()
        if t_next > now:
            time.sleep(t_next - now)
        n, d, k, seed = job_spec(i)
        with cv:
            cv.wait_for(lambda: free)
            slot = free.pop()
        t_arr = time.perf_counter()
        gen.generate_device(K.GEN_LOWRANK, n, d, d, seed, pool[slot].ptr)
        # lists of 100 while building: neighbours of neighbours of a 10-wide list see 100
        # candidates per point instead of 10^4, which does not reach recall 0.94
        prm = K.params(k=k, leaf_max=a.leaf, small_max=3000, iterations=a.iterations,
                       rounds=a.rounds, non_passes=a.non, seed=seed, search_k=max(k, 100))
        t = srv.submit(prm, device_ptr=pool[slot].ptr, n=n, d=d)
        th = threading.Thread(target=finish, args=(i, t, slot, n, d, k, seed, t_arr))
        th.start()
        threads.append(th)
        i += 1
        t_next += rng.exponential(1.0 / a.rate)
    for th in threads:
        th.join()
    t_end = max(r["done"] for r in results)
    srv.close()
    lat = np.array([(r["done"] - r["arrival"]) * 1e3 for r in sorted(results, key=lambda r: r["i"])])
    pts = sum(r["n"] for r in results)
    log(f"[bench] serving: {len(results)} jobs, {pts} points; latency p50 {np.percentile(lat, 50):.0f} "
        f"p99 {np.percentile(lat, 99):.0f} ms; checking recall")

    # recall@k of every job on its sample (exact FP64 bounded by the job's own lists)
    recalls = []
    chk = K.Builder(K.params(k=100))
    dX = pool[0]
    for r in sorted(results, key=lambda r: r["i"]):
        n, d, k = r["n"], r["d"], r["k"]
        gen.generate_device(K.GEN_LOWRANK, n, d, d, r["seed"], dX.ptr)
        nq = r["q"].size
        dq = K.DeviceArray(nq * 4)
        dq.upload(r["q"])
        dl = K.DeviceArray(nq * k * 4)
        dl.upload(r["got"])
        dub = K.DeviceArray(nq * 8)
        chk.set_params(K.params(k=k))
        chk.exact_upper_from_lists(dX.ptr, n, d, d, dq.ptr, nq, dl.ptr, k, dub.ptr)
        dti = K.DeviceArray(nq * k * 4)
        dtd = K.DeviceArray(nq * k * 8)
        chk.exact_knn_device(dX.ptr, n, d, d, dq.ptr, nq, k, dub.ptr, 4 * k + 256, dti.ptr, dtd.ptr)
        recalls.append(K.recall(r["got"], dti.download((nq, k), np.uint32)))
        log(f"[bench] job {r['i']}: n={n} d={d} k={k} latency {(r['done'] - r['arrival']) * 1e3:.0f} ms "
            f"recall@{k} {recalls[-1]:.4f}")
    recalls = np.array(recalls)
    out = {
        "metric": "serving: KNN-graph build jobs, Poisson arrivals; p99 job latency",
        "value": round(float(np.percentile(lat, 99)), 1), "unit": "ms (p99 job latency)",
        "higher_is_better": False, "n_gpus": 1, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"Poisson {a.rate} jobs/s for {a.duration} s; N log-uniform [1e4, 2e6], "
                               f"d in {{64,100,128}}, k in {{10,100}}, low-rank recipe per job seed",
                   "workers": a.workers, "budget_gb": a.budget_gb, "rounds": a.rounds,
                   "search_k": "max(k, 100)",
                   "non_passes": a.non, "leaf": a.leaf},
        "jobs": len(results), "points": int(pts),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 1),
                       "p95": round(float(np.percentile(lat, 95)), 1),
                       "p99": round(float(np.percentile(lat, 99)), 1),
                       "max": round(float(lat.max()), 1)},
        "aggregate_points_per_s": round(pts / (t_end - t0), 1),
        "recall_at_k": {"min": round(float(recalls.min()), 4), "mean": round(float(recalls.mean()), 4),
                        "jobs_below_0.94": int((recalls < 0.94).sum()), "sample": "10k points per job (all if N < 10k)"},
    }
    print(json.dumps(out), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the build does not "
              "shard across GPUs in this version", file=sys.stderr)
        sys.exit(2)
    import sigmod23ann_amd as K

    if a.workload == "zipf128":
        zipf_round(a, K)
        return
    if a.workload == "serving":
        serving(a, K)
        return

    if a.copy_calibration_mib > 0:
        copy_calibration(a.copy_calibration_mib)

    prec = {"fp32": K.PREC_FP32, "fp16": K.PREC_FP16, "bf16": K.PREC_BF16}[a.precision]
    flags = ((K.FLAG_KD_SPLIT if a.splitter == "kd" else 0) | (K.FLAG_AUTOTUNE if a.autotune else 0)
             | (K.FLAG_NON_SCREEN16 if a.non_screen16 else 0) | (K.FLAG_NON_FP32_GRAM if a.non_fp32_gram else 0)
             | leaf_flags(a, K))
    prm = K.params(k=a.k, leaf_max=a.leaf, small_max=3000, iterations=a.iterations,
                   rounds=a.rounds, non_passes=a.non, precision=prec, seed=a.seed, flags=flags)
    b = K.Builder(prm)
    n, d, k = a.n, a.d, a.k
    ld = (d + 3) & ~3
    log(f"[bench] generating {n}x{d} low-rank points on the device")
    dX = K.DeviceArray(n * ld * 4)
    b.generate_device(K.GEN_LOWRANK, n, d, ld, a.seed, dX.ptr)
    dIds = K.DeviceArray(n * k * 4)
    dD = K.DeviceArray(n * k * 4)

    def run_builds(builder, tag):
        for w in range(a.warmup):
            t0 = time.perf_counter()
            builder.build_device(dX.ptr, n, d, ld, dIds.ptr, dD.ptr)
            log(f"[bench] {tag} warmup {w}: {time.perf_counter() - t0:.3f} s")
        times, stats = [], []
        for s in range(a.steps):
            K.lib().knng_synchronize()
            t0 = time.perf_counter()
            builder.build_device(dX.ptr, n, d, ld, dIds.ptr, dD.ptr)
            K.lib().knng_synchronize()
            dt = time.perf_counter() - t0
            times.append(dt)
            st = builder.stats()
            stats.append(st)
            log(f"[bench] {tag} step {s}: {dt:.3f} s  split {st['split_ms']:.1f} ms "
                f"(side {st['side_ms']:.1f}) leaf {st['leaf_ms']:.1f} ms non {st['non_ms']:.1f} ms")
        return times, stats

    rng = np.random.default_rng(a.seed + 1)
    nq = min(a.recall_queries, n)
    q = np.sort(rng.choice(n, size=nq, replace=False)).astype(np.uint32) if nq > 0 else None

    def measure_recall(builder):
        t0 = time.perf_counter()
        ids = dIds.download((n, k), np.uint32)
        got = np.ascontiguousarray(ids[q])
        dq = K.DeviceArray(nq * 4)
        dl = K.DeviceArray(nq * k * 4)
        dub = K.DeviceArray(nq * 8)
        dti = K.DeviceArray(nq * k * 4)
        dtd = K.DeviceArray(nq * k * 8)
        truth = np.empty((nq, k), np.uint32)
        # batches of queries: a weak graph gives loose bounds, and the exact scan's re-scan of
        # queries with many candidates is memory-capped per call (halve a batch that overflows)
        q0, batch = 0, 100_000
        while q0 < nq:
            m = min(batch, nq - q0)
            dq.upload(q[q0:q0 + m])
            dl.upload(got[q0:q0 + m])
            builder.exact_upper_from_lists(dX.ptr, n, d, ld, dq.ptr, m, dl.ptr, k, dub.ptr)
            try:
                builder.exact_knn_device(dX.ptr, n, d, ld, dq.ptr, m, k, dub.ptr, 4 * k + 256,
                                         dti.ptr, dtd.ptr)
            except K.KnngError as e:
                if e.code != -5 or batch <= 64:
                    raise
                batch //= 2
                continue
            truth[q0:q0 + m] = dti.download((m, k), np.uint32)
            q0 += m
        rec = K.recall(got, truth)
        log(f"[bench] recall@{k} on {nq} sampled points vs FP64: {rec:.4f} "
            f"({time.perf_counter() - t0:.1f} s)")
        return rec, got

    # Reference forms at the same rounds and seed: the headline FP32 form (16-bit screens with
    # exact margins) when another precision is measured, and the plain FP32 GPU form (FP32 Gram in
    # both stages, no 16-bit Gram anywhere); both give the exact FP32 lists.
    cmp32 = None
    if a.compare_fp32 and nq > 0:
        cmp32 = {}
        forms = ([("fp32", 0)] if a.precision != "fp32" else []) + \
                [("plain_fp32", K.FLAG_LEAF_FP32_GRAM | K.FLAG_NON_FP32_GRAM)]
        for tag, fl in forms:
            b32 = K.Builder(K.params(k=a.k, leaf_max=a.leaf, small_max=3000, iterations=a.iterations,
                                     rounds=a.rounds, non_passes=a.non, precision=K.PREC_FP32,
                                     seed=a.seed, flags=fl))
            t32, _ = run_builds(b32, tag)
            r32, _ = measure_recall(b32)
            cmp32[f"{tag}_points_per_s"] = round(n / float(np.mean(t32)), 1)
            cmp32[f"{tag}_recall"] = round(r32, 4)
            del b32

    times, stats = run_builds(b, a.precision)
    mean_t = float(np.mean(times))
    value = n / mean_t
    st = stats[-1]

    # ---- the same build with its output copied to host memory (what knng_build's caller pays) --
    K.lib().knng_synchronize()
    t0 = time.perf_counter()
    b.build_device(dX.ptr, n, d, ld, dIds.ptr, dD.ptr)
    K.lib().knng_synchronize()
    t_b = time.perf_counter() - t0
    t1 = time.perf_counter()
    dIds.download((n, k), np.uint32)
    t_d2h = time.perf_counter() - t1
    to_host = {"points_per_s": round(n / (t_b + t_d2h), 1), "build_ms": round(t_b * 1e3, 2),
               "ids_d2h_ms": round(t_d2h * 1e3, 2), "ids_bytes": n * k * 4,
               "definition": "one more build from device-resident points, then its n*k ids "
                             "copied to pageable host memory (knng_memcpy_d2h: through two pinned "
                             "32 MB buffers, copied on by host threads); value excludes the copy"}
    log(f"[bench] to host: build {t_b * 1e3:.1f} ms + ids copy {t_d2h * 1e3:.1f} ms "
        f"-> {n / (t_b + t_d2h):.0f} points/s")

    # ---- recall against exact FP64 k-NN of sampled points ---------------------------------------
    rec, prec_block = None, None
    if nq > 0:
        rec, got = measure_recall(b)
        # returned distances against FP64 recomputation of the same pairs
        ne = min(a.error_queries, nq)
        if ne > 0:
            Xh = dX.download((n, ld), np.float32)
            ds_all = dD.download((n, k), np.float32)
            qe = q[:ne]
            ids_e = got[:ne].astype(np.int64)
            errs, rels = [], []
            for c0 in range(0, ne, 1000):
                sl = slice(c0, min(ne, c0 + 1000))
                xq = Xh[qe[sl], :d].astype(np.float64)
                xn = Xh[ids_e[sl], :d].astype(np.float64)
                d64 = ((xn - xq[:, None, :]) ** 2).sum(-1)
                d32 = ds_all[qe[sl]].astype(np.float64)
                m = d32 < 3e38
                errs.append(np.abs(d32 - d64)[m])
                rels.append((np.abs(d32 - d64) / np.maximum(d64, 1e-30))[m])
            err = np.concatenate(errs)
            relerr = np.concatenate(rels)
            del Xh, ds_all
            non_screen = ("fp32 Gram with exact error margin" if a.non_fp32_gram or d not in (61, 62, 63, 64, 97, 98, 99, 100, 125, 126, 127, 128)
                          else "16-bit Gram of rows centred on each q" + (" with exact error margin" if a.precision == "fp32" else ", no margin")
                          if a.non_screen16 else "fp16 Gram of the rows (scaled by powers of two unless every squared norm lies in [2^-12, 2^28]) with exact error margin")
            prec_block = {"precision": a.precision,
                          "screen": {"fp32": "leaf: fp16 Gram with exact error margin",
                                     "fp16": "leaf: fp16 Gram, no margin",
                                     "bf16": "leaf: bf16 Gram, no margin"}[a.precision]
                                    + "; neighbours of neighbours: " + non_screen,
                          "rerank": ("exact reference-order FP32 for every screen survivor (the leaf screen's "
                                     "bound lowered to each point's k-th smallest in-bucket upper bound)"
                                     if a.precision == "fp32" else
                                     "exact reference-order FP32 for the screen survivors among about the "
                                     "160 nearest by the screen per point and leaf bucket"),
                          "dist_abs_err_vs_fp64": {"max": float(err.max()), "mean": float(err.mean())},
                          "dist_rel_err_vs_fp64": {"max": float(relerr.max()),
                                                   "mean": float(relerr.mean())},
                          "error_queries": ne}
            if cmp32:
                prec_block.update(cmp32)
                for tag in ("fp32", "plain_fp32"):
                    if f"{tag}_points_per_s" in cmp32:
                        prec_block[f"recall_delta_vs_{tag}"] = round(rec - cmp32[f"{tag}_recall"], 4)
                        prec_block[f"speedup_vs_{tag}"] = round((n / mean_t) / cmp32[f"{tag}_points_per_s"], 3)
            log(f"[bench] distance error vs FP64 on {ne} points: max {err.max():.3e} "
                f"mean {err.mean():.3e} (relative max {relerr.max():.3e})")

    # ---- roofline of the dominant kernel -----------------------------------------------------
    # Algorithmic bytes/flops per launch over the HIP-event launch time:
    #   side test: 2d + 4 + 1 bytes per position (fp16 row, permutation entry, flag byte; 4d for
    #   the FP32-only form)
    #   leaf: d * b(b-1) flops per bucket of b points (unique pairs x 2d)
    #   NoN screen: 4dk bytes per q (rows of its list) + (4d + 4k) per reverse entry (row, snapshot)
    cands = []
    side_bytes = st["side_rows"] * side_row_bytes(a, d)     # row, perm entry, flag byte
    if st["side_ms"] > 0:
        cands.append(("k_side_count", "hbm", side_bytes, st["side_ms"], st["side_launches"]))
    leaf_flops = 2.0 * d * st["leaf_pairs"]                  # d * sum b(b-1): unique pairs x 2d
    if st["leaf_ms"] > 0:
        cands.append(("k_leaf", "mfma", leaf_flops, st["leaf_ms"], st["leaf_launches"]))
    if st["non_screen_ms"] > 0:
        cands.append(("k_non_reverse", "hbm", st["non_screen_bytes"], st["non_screen_ms"],
                      st["non_screen_launches"]))
    name, bound, work, ms, launches = max(cands, key=lambda c: c[3])
    if bound == "hbm":
        ach = work / (ms * 1e-3) / 1e9
        roof = {"kernel": name, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_launch": work / max(1, launches)}
    else:
        ach = work / (ms * 1e-3) / 1e12
        roof = {"kernel": name, "bound": "mfma", "achieved": round(ach, 2),
                "peak": MFMA_FP16_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / MFMA_FP16_PEAK_TFLOPS, 5), "traffic": None,
                "flops_per_launch": work / max(1, launches)}
    roof["traffic"] = measured_traffic(name, traffic_key(a))
    roof["launches"] = launches
    roof["avg_launch_ms"] = round(ms / max(1, launches), 4)
    roof["share_of_build"] = round(ms / st["total_ms"], 3) if st["total_ms"] > 0 else None
    roof["others"] = {c[0]: {"ms": round(c[3], 1),
                             ("GB/s" if c[1] == "hbm" else "TFLOP/s"):
                                 round(c[2] / (c[3] * 1e-3) / (1e9 if c[1] == "hbm" else 1e12), 2)}
                      for c in cands if c[0] != name}

    # ---- CPU baseline: the oracle's restatement on a prefix sample ---------------------------
    cpu = None
    if not a.no_cpu and a.cpu_sample > 0:
        import oracle
        ns = min(a.cpu_sample, n)
        Xh = dX.download((n, ld), np.float32)[:ns, :d].copy()
        # every core this process may run on, up to 16: the host's share for one GPU
        avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        threads = a.cpu_threads or min(avail, 16)
        op = oracle.make_params(k=k, leaf_max=a.leaf, small_max=3000, iterations=a.iterations,
                                rounds=a.rounds, non_passes=a.non, seed=a.seed,
                                splitter=1 if a.splitter == "kd" else 0)
        t0 = time.perf_counter()
        _, _, _, ost = oracle.build_graph(Xh, op, threads=threads)
        ct = time.perf_counter() - t0
        # Extrapolation to the full n: leaf, NoN and finalize work grows with the point count;
        # the split grows with the point count times the tree depth (levels ~ log2(n / leaf)).
        depth = max(1.0, np.log2(n / a.leaf)) / max(1.0, np.log2(ns / a.leaf))
        t_full = (ost.split_s * depth + ost.leaf_s + ost.non_s + ost.finalize_s) * (n / ns)
        cpu = {"value": round(ns / ct, 1), "unit": "points/s", "cores": threads, "kind": "port",
               "sample": f"first {ns} points of the same data, same rounds/leaf/k/NoN, "
                         f"{ct:.1f} s on {threads} threads (of {avail} available; oracle built "
                         f"-O3 -march=x86-64-v3)",
               "stages_s": {"split": round(ost.split_s, 2), "leaf": round(ost.leaf_s, 2),
                            "non": round(ost.non_s, 2), "finalize": round(ost.finalize_s, 2)},
               "extrapolated_full_n": {"seconds": round(t_full, 1),
                                       "points_per_s": round(n / t_full, 1),
                                       "rule": "stage times x n/sample, split also x "
                                               "log2(n/leaf)/log2(sample/leaf)"}}
        log(f"[bench] cpu baseline: {ns / ct:.0f} points/s ({ct:.1f} s, {threads} threads); "
            f"extrapolated to {n}: {t_full:.0f} s")

    lat = sorted(t * 1e3 for t in times)
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "points/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(mean_t * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": a.precision if a.precision == "fp32" else f"{a.precision} screen, fp32 re-rank",
        "data": "synthetic",
        "config": {"workload": f"low-rank {n}x{d} (24 latent, unit norm), k={k}, leaf<={a.leaf}",
                   "n": n, "d": d, "k": k, "rounds": a.rounds, "iterations": a.iterations,
                   "non_passes": a.non, "precision": a.precision, "splitter": a.splitter, "autotune": a.autotune,
                   "leaf_mode": a.leaf_mode, "bucket_limit": not a.no_bucket_limit,
                   "side_test": "fp32" if a.side_fp32 or a.splitter == "kd" else "fp16 with exact bound",
                   "recall_at_k": None if rec is None else round(rec, 4),
                   "recall_queries": nq, "parallelism": "single GPU"},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2),
                       "p99": round(float(np.percentile(lat, 99)), 2)},
        "stages_ms": {"split": round(st["split_ms"], 2), "side": round(st["side_ms"], 2),
                      "leaf": round(st["leaf_ms"], 2), "non": round(st["non_ms"], 2),
                      "total": round(st["total_ms"], 2)},
        "non_overflow_points": int(st["non_overflow"]),
        "to_host": to_host,
        "precision_form": prec_block,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
