"""The benchmark configurations at their own sizes, through the C ABI:

- C0: 100k x 100 Gaussian mixture (500 centres), k = 100, 2 Lloyd iterations, leaves <= 256,
  8 rounds, 1 NoN pass: lists equal to the oracle's up to ties, recall@100 against exact FP64 on
  every point.
- C1/C2: 10M x 100 low-rank points, k = 100, leaves <= 400, 6 rounds, 2 NoN passes, in the exact
  FP32 form and the BF16 / FP16 screening forms: properties that hold at any size, recall@100
  against FP64, and the recall delta of each screening form against the FP32 form.
- C3: the Zipf(1.1) mixture generator against its recipe; one round (split + all buckets, leaves
  <= 4096) at 300k x 128 against the oracle, and at 10M x 128 through per-leaf oracle checks.
"""
import numpy as np
import pytest

import oracle
import sigmod23ann_amd as K

pytestmark = pytest.mark.gpu

F32MAX = np.finfo(np.float32).max


def _exact_truth(b, dX, n, d, ld, q, lists, k):
    """FP64 exact k-NN of queries q on the device (pinned against the oracle by
    test_exact_knn_gpu_matches_oracle), bounded by each query's own list."""
    nq = q.size
    dq = K.DeviceArray(nq * 4)
    dq.upload(q)
    dl = K.DeviceArray(nq * k * 4)
    dl.upload(np.ascontiguousarray(lists, np.uint32))
    dub = K.DeviceArray(nq * 8)
    b.exact_upper_from_lists(dX.ptr, n, d, ld, dq.ptr, nq, dl.ptr, k, dub.ptr)
    dti = K.DeviceArray(nq * k * 4)
    dtd = K.DeviceArray(nq * k * 8)
    b.exact_knn_device(dX.ptr, n, d, ld, dq.ptr, nq, k, dub.ptr, 4 * k + 256, dti.ptr, dtd.ptr)
    return dti.download((nq, k), np.uint32)


def _assert_ties_only(X, ids, ds, o_ids, o_ds):
    np.testing.assert_array_equal(ds, o_ds)
    boundary = ds == ds[:, -1:]
    bad = (ids != o_ids) & ~boundary
    assert not bad.any(), f"{int(bad.sum())} ids differ away from ties"
    for r in np.where((ids != o_ids).any(1))[0]:
        valid = ds[r] < F32MAX
        row = ids[r][valid]
        assert len(set(row.tolist())) == row.size and r not in row
        for j in ids[r][boundary[r] & valid]:
            assert oracle.dist(X[r], X[j]) == ds[r, -1]


# ---------------------------------------------------------------------------------------------
# C0
# ---------------------------------------------------------------------------------------------
@pytest.mark.timeout(300)
def test_c0_full_size_matches_oracle_and_fp64_recall():
    n, d, k = 100_000, 100, 100
    rng = np.random.default_rng(61)
    c = rng.uniform(-10, 10, size=(500, d)).astype(np.float32)
    X = (c[rng.integers(0, 500, n)] + rng.normal(size=(n, d))).astype(np.float32)
    kw = dict(k=k, leaf_max=256, small_max=3000, iterations=2, rounds=8, non_passes=1, seed=1)
    b = K.Builder(K.params(**kw))
    ids, ds = b.build(X, return_dists=True)
    o_ids, o_ds, _, _ = oracle.build_graph(X, oracle.make_params(**kw))
    _assert_ties_only(X, ids, ds, o_ids, o_ds)
    dX = K.DeviceArray(X.nbytes)
    dX.upload(X)
    q = np.arange(n, dtype=np.uint32)
    truth = _exact_truth(b, dX, n, d, d, q, ids, k)
    rec = K.recall(ids, truth)
    print(f"C0 recall@100 on all {n} points: {rec:.4f}")
    assert rec >= 0.94, rec


# ---------------------------------------------------------------------------------------------
# C1 / C2 at 10M
# ---------------------------------------------------------------------------------------------
def _full_size_properties(X, ids, ds, rng):
    n = ids.shape[0]
    step = 1_000_000
    for r0 in range(0, n, step):
        i, s = ids[r0:r0 + step], ds[r0:r0 + step]
        assert np.isfinite(s).all() and (i < n).all()
        assert (s[:, 1:] >= s[:, :-1]).all()
        tie = s[:, 1:] == s[:, :-1]
        assert (i[:, 1:][tie] > i[:, :-1][tie]).all()
        srt = np.sort(i, axis=1)
        assert (srt[:, 1:] != srt[:, :-1]).all()
        assert (i != np.arange(r0, r0 + i.shape[0], dtype=np.uint32)[:, None]).all()
    for r in rng.choice(n, 200, replace=False):
        got = np.array([oracle.dist(X[r], X[j]) for j in ids[r]], np.float32)
        np.testing.assert_array_equal(ds[r], got)


@pytest.mark.timeout(900)
def test_c1_c2_full_size_forms():
    # One builder, the same device points and the same queries for the three forms. Every form
    # re-ranks its survivors in exact FP32, so returned distances are bit-equal to the reference
    # order's; the screening forms may only lose candidates, which the recall delta measures.
    n, d, k = 10_000_000, 100, 100
    kw = dict(k=k, leaf_max=400, small_max=3000, iterations=1, rounds=6, non_passes=2, seed=2023)
    b = K.Builder(K.params(**kw))
    dX = K.DeviceArray(n * d * 4)
    b.generate_device(K.GEN_LOWRANK, n, d, d, 2023, dX.ptr)
    X = dX.download((n, d), np.float32)
    dI = K.DeviceArray(n * k * 4)
    dD = K.DeviceArray(n * k * 4)
    rng = np.random.default_rng(3)
    q = np.sort(rng.choice(n, 4000, replace=False)).astype(np.uint32)
    recalls = {}
    for name, prec in [("fp32", K.PREC_FP32), ("bf16", K.PREC_BF16), ("fp16", K.PREC_FP16)]:
        b.set_params(K.params(precision=prec, **kw))
        b.build_device(dX.ptr, n, d, d, dI.ptr, dD.ptr)
        ids = dI.download((n, k), np.uint32)
        ds = dD.download((n, k), np.float32)
        _full_size_properties(X, ids, ds, np.random.default_rng(5))
        truth = _exact_truth(b, dX, n, d, d, q, ids[q], k)
        recalls[name] = K.recall(ids[q], truth)
        print(f"{name}: recall@100 {recalls[name]:.4f} on {q.size} points, build "
              f"{b.stats()['total_ms']:.0f} ms")
        del ids, ds
        assert recalls[name] >= 0.94, (name, recalls[name])
    for name in ("bf16", "fp16"):
        assert recalls[name] >= recalls["fp32"] - 0.002, (name, recalls)


# ---------------------------------------------------------------------------------------------
# C3: Zipf mixture
# ---------------------------------------------------------------------------------------------
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix64(z):
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _zipf_recipe(seed, n, d, nc=20000):
    """The Zipf workload's recipe, restated: centres uniform in [-10, 10]^d, weights (c+1)^-1.1,
    sigma_c uniform in [0.3, 2.0] (host draws from one splitmix64 stream of the seed), and each
    point's cluster from its own counter draw against the cumulative float weights."""
    s0 = int(_mix64(np.uint64(seed ^ 0xA0761D6478BD642F)))
    m = nc * d + nc
    with np.errstate(over="ignore"):
        st = np.uint64(s0) + np.uint64(0x9E3779B97F4A7C15) * np.arange(1, m + 1, dtype=np.uint64)
    u = (_mix64(st) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    C = (-10.0 + 20.0 * u[: nc * d]).astype(np.float32).reshape(nc, d)
    sg = (0.3 + 1.7 * u[nc * d:]).astype(np.float32)
    w = 1.0 / np.power(np.arange(1, nc + 1, dtype=np.float64), 1.1)
    run = np.cumsum(w)                       # sequential sums, as the host adds them
    cdf = (run / run[-1]).astype(np.float32)
    cdf[-1] = 1.0
    with np.errstate(over="ignore"):
        hp = _mix64(np.uint64(seed) ^ _mix64(np.arange(n, dtype=np.uint64) + np.uint64(0x5851F42D4C957F2D)))
    up = ((hp >> np.uint64(40)).astype(np.float32) + np.float32(1.0)) * np.float32(1.0 / 16777216.0)
    lab = np.searchsorted(cdf, up, side="left")
    return C, sg, w / w.sum(), lab


def test_zipf_generator_matches_its_recipe():
    n, d, seed = 100_000, 128, 77
    b = K.Builder(K.params(k=10))
    dX = K.DeviceArray(n * d * 4)
    b.generate_device(K.GEN_ZIPF, n, d, d, seed, dX.ptr)
    X = dX.download((n, d), np.float32)
    C, sg, p, lab = _zipf_recipe(seed, n, d)
    # every point lies nearest to the centre its draw names (centres ~90 apart, noise <= 2 per dim)
    cn = (C.astype(np.float64) ** 2).sum(1)
    for r0 in range(0, n, 10_000):
        Y = X[r0:r0 + 10_000].astype(np.float64)
        near = np.argmin(cn[None, :] - 2.0 * Y @ C.T.astype(np.float64), axis=1)
        np.testing.assert_array_equal(near, lab[r0:r0 + 10_000])
    # Zipf(1.1) weights: the heaviest clusters hold n p_c points within 5 binomial sigma
    cnt = np.bincount(lab, minlength=C.shape[0])
    e = n * p[:50]
    assert (np.abs(cnt[:50] - e) <= 5 * np.sqrt(e * (1 - p[:50])) + 1).all()
    assert cnt[0] > 0.1 * n
    # per-cluster noise: mean squared deviation per dimension ~ sigma_c^2
    dev = ((X.astype(np.float64) - C[lab]) ** 2).sum(1) / d
    for c in np.where(cnt >= 200)[0]:
        v = dev[lab == c].mean()
        assert abs(v / float(sg[c]) ** 2 - 1.0) < 0.05, (c, v, sg[c])
    assert 0.3 <= sg.min() and sg.max() <= 2.0


@pytest.mark.timeout(300)
def test_c3_zipf_round_matches_oracle():
    # One round of the batching form at 300k x 128, leaves <= 4096 (the big-bucket kernel for
    # leaves above 512): split permutation and leaves bit-exact, lists equal up to ties.
    n, d, k = 300_000, 128, 100
    b = K.Builder(K.params(k=k, leaf_max=4096, small_max=3000, iterations=1, rounds=1,
                           non_passes=0, seed=9))
    dX = K.DeviceArray(n * d * 4)
    b.generate_device(K.GEN_ZIPF, n, d, d, 9, dX.ptr)
    X = dX.download((n, d), np.float32)
    del dX
    p = oracle.make_params(k=k, leaf_max=4096, small_max=3000, iterations=1, rounds=1, seed=9)
    perm, ls, lz = b.split_round(X, 0)
    o_perm, o_ls, o_lz = oracle.split_round(X, p, 0)
    np.testing.assert_array_equal(perm, o_perm)
    np.testing.assert_array_equal(ls, o_ls)
    np.testing.assert_array_equal(lz, o_lz)
    assert lz.max() > 512 and lz.min() < 512
    print(f"{lz.size} buckets, sizes {lz.min()}..{lz.max()}")
    ids, ds = b.leaf_stage(X, perm, ls, lz)
    o_ids, o_ds, _ = oracle.leaf_stage(X, k, perm, ls, lz)
    _assert_ties_only(X, ids, ds, o_ids, o_ds)


@pytest.mark.timeout(600)
def test_c3_zipf_round_full_size_per_leaf_oracle():
    # The C3 round itself (10M x 128, leaves <= 4096): every list holds only ids of the point's own
    # bucket, and for sampled buckets (the largest among them) the lists equal the oracle's
    # all-pairs stage over that bucket, up to ties.
    n, d, k = 10_000_000, 128, 100
    b = K.Builder(K.params(k=k, leaf_max=4096, small_max=3000, iterations=1, rounds=1,
                           non_passes=0, seed=2023))
    dX = K.DeviceArray(n * d * 4)
    b.generate_device(K.GEN_ZIPF, n, d, d, 2023, dX.ptr)
    X = dX.download((n, d), np.float32)
    del dX
    perm, ls, lz = b.split_round(X, 0)
    assert np.array_equal(np.sort(perm), np.arange(n, dtype=np.uint32))
    assert ls[0] == 0 and (ls[1:] == ls[:-1] + lz[:-1]).all() and ls[-1] + lz[-1] == n
    assert lz.max() <= 4096
    print(f"{lz.size} buckets, sizes {lz.min()}..{lz.max()}")
    ids, ds = b.leaf_stage(X, perm, ls, lz)
    leaf_of = np.empty(n, np.int64)
    leaf_of[perm] = np.repeat(np.arange(lz.size), lz)
    valid = ds < F32MAX
    for r0 in range(0, n, 1_000_000):
        sl = slice(r0, r0 + 1_000_000)
        own = leaf_of[sl][:, None]
        assert ((leaf_of[np.where(valid[sl], ids[sl], 0)] == own) | ~valid[sl]).all()
    rng = np.random.default_rng(4)
    pick = set(rng.choice(lz.size, 12, replace=False).tolist()) | {int(np.argmax(lz))}
    for li in sorted(pick):
        # the oracle's all-pairs stage over this bucket alone, on its rows in bucket order
        rows = perm[ls[li]:ls[li] + lz[li]]
        b_ = rows.size
        o_loc, o_ds, _ = oracle.leaf_stage(X[rows], k, np.arange(b_, dtype=np.uint32),
                                           np.zeros(1, np.uint32), np.array([b_], np.uint32))
        o_ids = np.where(o_ds < F32MAX, rows[np.minimum(o_loc, b_ - 1)], 0xFFFFFFFF).astype(np.uint32)
        _assert_ties_only_rows(X, rows, ids[rows], ds[rows], o_ids, o_ds)


def _assert_ties_only_rows(X, rows, i, s, oi, od):
    np.testing.assert_array_equal(s, od)
    boundary = s == s[:, -1:]
    assert not ((i != oi) & ~boundary).any()
    for t in np.where((i != oi).any(1))[0]:
        r = rows[t]
        for j in i[t][boundary[t] & (s[t] < F32MAX)]:
            assert oracle.dist(X[r], X[j]) == s[t, -1]
