"""GPU parity: every stage through the C ABI against the CPU oracle on the same seeded inputs.

Integer/index work (permutations, leaf ranges, lo sizes) must match bit for bit; top-k lists must
match exactly in distance (both sides compute the reference-order FP32 distance) and in ids except
where distances tie at the k-th place (the reference keeps the first offered, the GPU the smaller
id). Sizes keep the oracle to a few seconds.
"""
import numpy as np
import pytest

import oracle
import sigmod23ann_amd as K

pytestmark = pytest.mark.gpu


def mixture(n, d, centres, seed, spread=10.0, sigma=1.0):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, size=(centres, d)).astype(np.float32)
    lab = rng.integers(0, centres, n)
    return (c[lab] + sigma * rng.normal(size=(n, d))).astype(np.float32)


def lowrank(n, d, seed, latent=24):
    rng = np.random.default_rng(seed)
    W = rng.normal(size=(latent, d))
    z = rng.normal(size=(n, latent)) / np.sqrt(1.0 + np.arange(latent))
    X = z @ W + 0.05 * rng.normal(size=(n, d))
    return (X / np.linalg.norm(X, axis=1, keepdims=True)).astype(np.float32)


def assert_lists_equal_up_to_ties(X, ids, ds, o_ids, o_ds):
    """Same sorted distances; same ids wherever the distance is not tied with the k-th. In a row
    whose ids differ at the tied k-th distance, the row's ids are distinct, exclude the point
    itself, and each tied id lies at exactly the tied distance (reference-order distance)."""
    ids = np.asarray(ids)
    o_ids = np.asarray(o_ids)
    np.testing.assert_array_equal(ds, o_ds)
    boundary = ds == ds[:, -1:]
    bad = (ids != o_ids) & ~boundary
    assert not bad.any(), f"{int(bad.sum())} ids differ away from ties, first rows {np.where(bad.any(1))[0][:5]}"
    fmax = np.finfo(np.float32).max
    for r in np.where((ids != o_ids).any(1))[0]:
        valid = ds[r] < fmax
        row = ids[r][valid]
        assert len(set(row.tolist())) == row.size, f"row {r}: repeated id"
        assert r not in row, f"row {r}: lists itself"
        for j in ids[r][boundary[r] & valid]:
            assert j < X.shape[0] and oracle.dist(X[r], X[j]) == ds[r, -1], f"row {r}: id {j} not at the tied distance"


def assert_valid_lists(X, ids, ds):
    n, k = ids.shape
    for r in range(0, n, max(1, n // 200)):
        row = ids[r][ds[r] < np.finfo(np.float32).max]
        assert r not in row
        assert len(set(row.tolist())) == row.size
        assert np.all(np.diff(ds[r]) >= 0)
        for j, dv in zip(row[:5], ds[r][:5]):
            assert oracle.dist(X[r], X[j]) == dv


# ---------------------------------------------------------------------------------------------
# Split
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("d", [100, 64, 128, 37])
def test_partition_injected_bitexact(d):
    X = mixture(20000, d, 50, 1)
    rng = np.random.default_rng(2)
    perm0 = rng.permutation(20000).astype(np.uint32)
    starts = np.array([0, 5000, 5001, 9000, 15000], np.uint32)
    sizes = np.array([5000, 1, 3999, 6000, 5000], np.uint32)
    coefs = rng.normal(size=(5, d)).astype(np.float32)
    offs = (rng.normal(size=5) * 5).astype(np.float32)
    offs[1] = 1e30          # single-element node, every point on the hi side: unchanged
    b = K.Builder(K.params(k=10))
    perm, lo = b.partition_injected(X, perm0, starts, sizes, coefs, offs)
    o_perm, o_lo = oracle.partition_injected(X, perm0, starts, sizes, coefs, offs)
    np.testing.assert_array_equal(lo, o_lo)
    np.testing.assert_array_equal(perm, o_perm)


@pytest.mark.parametrize("d", [100, 37, 256])
def test_partition_injected_fp16_side_test_defers_to_fp32_rows(d):
    # The default side test reads fp16 rows and decides only outside its error bound; points on or
    # near the hyperplane, values that overflow fp16 and values in its subnormal range must reach
    # the FP32 rows' decision: the same flags as the oracle and as the FP32-only side test.
    rng = np.random.default_rng(4)
    n = 12000
    X = rng.integers(-3, 4, size=(n, d)).astype(np.float32)      # integer points: exact ties
    X[3000:6000] += rng.choice([-1e-5, 0.0, 1e-5], size=(3000, d)).astype(np.float32)
    X[6000:6100, 0] = 7e4                                          # beyond fp16's range
    X[6100:9000] *= 1e-6                                           # fp16 subnormal range
    perm0 = rng.permutation(n).astype(np.uint32)
    starts = np.array([0, 3000, 6000, 9000], np.uint32)
    sizes = np.array([3000, 3000, 3000, 3000], np.uint32)
    coefs = rng.integers(-2, 3, size=(4, d)).astype(np.float32)
    coefs[2] *= 1e-3
    offs = np.array([0.0, 1.0, 0.0, 2.0], np.float32)
    res = []
    for flags in (0, K.FLAG_SIDE_FP32):
        b = K.Builder(K.params(k=10, flags=flags))
        res.append(b.partition_injected(X, perm0, starts, sizes, coefs, offs))
    o_perm, o_lo = oracle.partition_injected(X, perm0, starts, sizes, coefs, offs)
    for perm, lo in res:
        np.testing.assert_array_equal(lo, o_lo)
        np.testing.assert_array_equal(perm, o_perm)
    # the integer nodes hold points exactly on their hyperplanes
    on = (X[perm0[:3000]].astype(np.float64) @ coefs[0].astype(np.float64) == offs[0]).sum()
    assert on > 0


def test_split_round_fp16_side_test_equals_fp32_side_test():
    X = lowrank(30000, 100, 8)
    kw = dict(k=10, leaf_max=200, small_max=3000, iterations=2, seed=8)
    out = [K.Builder(K.params(flags=f, **kw)).split_round(X, 1) for f in (0, K.FLAG_SIDE_FP32)]
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("n,d,leaf,small,iters,seed", [
    (50000, 100, 256, 3000, 2, 11),
    (30000, 100, 400, 3000, 1, 12),
    (20000, 64, 100, 500, 1, 13),
    (20000, 128, 256, 3000, 3, 14),
    (7000, 37, 50, 300, 2, 15),
])
def test_split_round_bitexact(n, d, leaf, small, iters, seed):
    X = mixture(n, d, 80, seed)
    kw = dict(k=10, leaf_max=leaf, small_max=small, iterations=iters, seed=seed)
    b = K.Builder(K.params(**kw))
    o_prm = oracle.make_params(**kw)
    for r in (0, 1):
        perm, ls, lz = b.split_round(X, r)
        o_perm, o_ls, o_lz = oracle.split_round(X, o_prm, r)
        np.testing.assert_array_equal(ls, o_ls)
        np.testing.assert_array_equal(lz, o_lz)
        np.testing.assert_array_equal(perm, o_perm)


def test_split_round_lowrank_and_alternating_leaf():
    X = lowrank(40000, 100, 3)
    kw = dict(k=10, leaf_max=199, leaf_max_odd=399, small_max=3000, iterations=1, seed=5)
    b = K.Builder(K.params(**kw))
    for r in (0, 1):
        perm, ls, lz = b.split_round(X, r)
        o_perm, o_ls, o_lz = oracle.split_round(X, oracle.make_params(**kw), r)
        np.testing.assert_array_equal(perm, o_perm)
        np.testing.assert_array_equal(lz, o_lz)
        assert lz.max() <= (199 if r == 0 else 399)


@pytest.mark.parametrize("n,d,leaf,seed", [(30000, 100, 399, 5), (12000, 37, 150, 6), (20000, 64, 255, 7)])
def test_kd_split_round_bitexact(n, d, leaf, seed):
    # random-rotation KD splitter (SolutionRandomKD): same directions, FP64 node statistics,
    # dimension draws and stable partition as the oracle
    X = mixture(n, d, 60, seed)
    kw = dict(k=10, leaf_max=leaf, seed=seed)
    b = K.Builder(K.params(flags=K.FLAG_KD_SPLIT, **kw))
    o_prm = oracle.make_params(splitter=1, **kw)
    for r in (0, 1):
        perm, ls, lz = b.split_round(X, r)
        o_perm, o_ls, o_lz = oracle.split_round(X, o_prm, r)
        np.testing.assert_array_equal(ls, o_ls)
        np.testing.assert_array_equal(lz, o_lz)
        np.testing.assert_array_equal(perm, o_perm)


def test_kd_split_identical_points():
    X = np.ones((3000, 16), np.float32)
    kw = dict(k=10, leaf_max=99, seed=3)
    perm, ls, lz = K.Builder(K.params(flags=K.FLAG_KD_SPLIT, **kw)).split_round(X, 0)
    o_perm, o_ls, o_lz = oracle.split_round(X, oracle.make_params(splitter=1, **kw), 0)
    np.testing.assert_array_equal(perm, o_perm)
    np.testing.assert_array_equal(lz, o_lz)


def test_split_degenerate_points():
    X = np.ones((5000, 16), np.float32)
    X[:10] += np.arange(10)[:, None]          # a few distinct rows, the rest identical
    kw = dict(k=10, leaf_max=100, small_max=1000, iterations=1, seed=3)
    perm, ls, lz = K.Builder(K.params(**kw)).split_round(X, 0)
    o_perm, o_ls, o_lz = oracle.split_round(X, oracle.make_params(**kw), 0)
    np.testing.assert_array_equal(perm, o_perm)
    np.testing.assert_array_equal(lz, o_lz)


# ---------------------------------------------------------------------------------------------
# Leaf stage, NoN, whole build
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("k,leaf,d", [(100, 256, 100), (10, 400, 100), (32, 128, 64), (100, 400, 128)])
def test_leaf_stage_matches_oracle(k, leaf, d):
    X = mixture(12000, d, 40, 21)
    kw = dict(k=k, leaf_max=leaf, small_max=3000, iterations=1, seed=9)
    o_prm = oracle.make_params(**kw)
    b = K.Builder(K.params(**kw))
    perm, ls, lz = oracle.split_round(X, o_prm, 0)
    ids, ds = b.leaf_stage(X, perm, ls, lz)
    o_ids, o_ds, o_sz = oracle.leaf_stage(X, k, perm, ls, lz)
    assert_lists_equal_up_to_ties(X, ids, ds, o_ids, o_ds)
    # second round merges into the existing sets
    perm2, ls2, lz2 = oracle.split_round(X, o_prm, 1)
    ids2, ds2 = b.leaf_stage(X, perm2, ls2, lz2, state=(ids, ds))
    o2 = oracle.leaf_stage(X, k, perm2, ls2, lz2, state=(o_ids, o_ds, o_sz))
    assert_lists_equal_up_to_ties(X, ids2, ds2, o2[0], o2[1])


@pytest.mark.parametrize("k,leaf,d", [(100, 256, 100), (10, 400, 100), (32, 128, 64), (100, 400, 128)])
def test_leaf_stage_big_kernel_matches_oracle(k, leaf, d):
    # the big-bucket kernel (64-row blocks, streamed column blocks) forced onto ordinary leaves
    X = mixture(12000, d, 40, 21)
    kw = dict(k=k, leaf_max=leaf, small_max=3000, iterations=1, seed=9)
    o_prm = oracle.make_params(**kw)
    b = K.Builder(K.params(flags=K.FLAG_BIG_LEAF, **kw))
    perm, ls, lz = oracle.split_round(X, o_prm, 0)
    ids, ds = b.leaf_stage(X, perm, ls, lz)
    o_ids, o_ds, o_sz = oracle.leaf_stage(X, k, perm, ls, lz)
    assert_lists_equal_up_to_ties(X, ids, ds, o_ids, o_ds)
    perm2, ls2, lz2 = oracle.split_round(X, o_prm, 1)
    ids2, ds2 = b.leaf_stage(X, perm2, ls2, lz2, state=(ids, ds))
    o2 = oracle.leaf_stage(X, k, perm2, ls2, lz2, state=(o_ids, o_ds, o_sz))
    assert_lists_equal_up_to_ties(X, ids2, ds2, o2[0], o2[1])


@pytest.mark.parametrize("k,leaf,d", [(100, 256, 100), (10, 400, 100), (32, 128, 64), (100, 400, 128),
                                      (20, 1500, 37)])
def test_leaf_fp32_gram_matches_oracle(k, leaf, d):
    # the FP32 Gram form of the leaf screen (16x16x4 MFMA over the FP32 rows, FP32 margin), every
    # bucket on the big-bucket kernel: the same lists as the oracle, on empty lists and merged
    X = mixture(12000, d, 40, 27)
    kw = dict(k=k, leaf_max=leaf, small_max=3000, iterations=1, seed=10)
    o_prm = oracle.make_params(**kw)
    b = K.Builder(K.params(flags=K.FLAG_LEAF_FP32_GRAM, **kw))
    perm, ls, lz = oracle.split_round(X, o_prm, 0)
    ids, ds = b.leaf_stage(X, perm, ls, lz)
    o_ids, o_ds, o_sz = oracle.leaf_stage(X, k, perm, ls, lz)
    assert_lists_equal_up_to_ties(X, ids, ds, o_ids, o_ds)
    perm2, ls2, lz2 = oracle.split_round(X, o_prm, 1)
    ids2, ds2 = b.leaf_stage(X, perm2, ls2, lz2, state=(ids, ds))
    o2 = oracle.leaf_stage(X, k, perm2, ls2, lz2, state=(o_ids, o_ds, o_sz))
    assert_lists_equal_up_to_ties(X, ids2, ds2, o2[0], o2[1])


def test_non_one_candidate_set_equals_two_sets():
    # one set of candidate buffers (each sub-pass merged before the next screen, as a build near
    # the device's capacity runs) against the default two sets: the same lists bit for bit
    X = lowrank(30000, 100, 19)
    kw = dict(k=100, leaf_max=400, small_max=3000, iterations=1, rounds=3, non_passes=2, seed=8)
    ids, ds = K.Builder(K.params(**kw)).build(X, return_dists=True)
    o_ids, o_ds = K.Builder(K.params(flags=K.FLAG_NON_ONE_SET, **kw)).build(X, return_dists=True)
    np.testing.assert_array_equal(o_ds, ds)
    np.testing.assert_array_equal(o_ids, ids)


def test_plain_fp32_form_equals_the_headline_form():
    # FP32 Gram in both stages (no 16-bit Gram anywhere) against the fp16-screened headline form:
    # both exact, so the same lists bit for bit
    X = lowrank(30000, 100, 17)
    kw = dict(k=100, leaf_max=400, small_max=3000, iterations=1, rounds=3, non_passes=1, seed=5)
    ids, ds = K.Builder(K.params(**kw)).build(X, return_dists=True)
    plain = K.Builder(K.params(flags=K.FLAG_LEAF_FP32_GRAM | K.FLAG_NON_FP32_GRAM, **kw))
    p_ids, p_ds = plain.build(X, return_dists=True)
    np.testing.assert_array_equal(p_ds, ds)
    np.testing.assert_array_equal(p_ids, ids)
    with pytest.raises(K.KnngError):
        K.Builder(K.params(precision=K.PREC_BF16, flags=K.FLAG_LEAF_FP32_GRAM, **kw))


@pytest.mark.parametrize("d,leaf", [(100, 2000), (128, 1500), (37, 3000)])
def test_leaf_stage_large_buckets(d, leaf):
    # buckets beyond the LDS kernel (up to `leaf` points) against brute force
    X = mixture(10000, d, 6, 23)
    kw = dict(k=100, leaf_max=leaf, small_max=3000, iterations=1, seed=5)
    o_prm = oracle.make_params(**kw)
    perm, ls, lz = oracle.split_round(X, o_prm, 0)
    assert lz.max() > 512
    b = K.Builder(K.params(**kw))
    ids, ds = b.leaf_stage(X, perm, ls, lz)
    o_ids, o_ds, o_sz = oracle.leaf_stage(X, 100, perm, ls, lz)
    assert_lists_equal_up_to_ties(X, ids, ds, o_ids, o_ds)
    perm2, ls2, lz2 = oracle.split_round(X, o_prm, 1)
    ids2, ds2 = b.leaf_stage(X, perm2, ls2, lz2, state=(ids, ds))
    o2 = oracle.leaf_stage(X, 100, perm2, ls2, lz2, state=(o_ids, o_ds, o_sz))
    assert_lists_equal_up_to_ties(X, ids2, ds2, o2[0], o2[1])


def test_build_large_leaves_matches_oracle():
    X = mixture(12000, 100, 30, 43)
    kw = dict(k=50, leaf_max=1500, small_max=1000, iterations=1, rounds=2, non_passes=1, seed=19)
    ids, ds = K.Builder(K.params(**kw)).build(X, return_dists=True)
    o_ids, o_ds, _, _ = oracle.build_graph(X, oracle.make_params(**kw))
    assert_lists_equal_up_to_ties(X, ids, ds, o_ids, o_ds)


def test_leaf_stage_lowrank_unit_norm():
    # close pairs on the unit sphere: the screen's margin must keep every true candidate
    X = lowrank(20000, 100, 8)
    kw = dict(k=100, leaf_max=400, small_max=3000, iterations=1, seed=4)
    o_prm = oracle.make_params(**kw)
    perm, ls, lz = oracle.split_round(X, o_prm, 0)
    ids, ds = K.Builder(K.params(**kw)).leaf_stage(X, perm, ls, lz)
    o_ids, o_ds, _ = oracle.leaf_stage(X, 100, perm, ls, lz)
    assert_lists_equal_up_to_ties(X, ids, ds, o_ids, o_ds)


def test_leaf_stage_autotune_matches_default_shape():
    # with KNNG_FLAG_AUTOTUNE the first launch of each bin above 128 points also times both
    # workgroup shapes and keeps the faster; the lists must equal the default shape's, on the
    # tuning call and on later calls that use the kept shape
    X = lowrank(20000, 100, 8)
    kw = dict(k=100, leaf_max=400, small_max=3000, iterations=1, seed=4)
    perm, ls, lz = oracle.split_round(X, oracle.make_params(**kw), 0)
    assert lz.max() > 256 and ((lz > 128) & (lz <= 256)).any()  # both tuned bins are launched
    ref_ids, ref_ds = K.Builder(K.params(**kw)).leaf_stage(X, perm, ls, lz)
    tuned = K.Builder(K.params(flags=K.FLAG_AUTOTUNE, **kw))
    for _ in range(2):
        ids, ds = tuned.leaf_stage(X, perm, ls, lz)
        np.testing.assert_array_equal(ids, ref_ids)
        np.testing.assert_array_equal(ds, ref_ds)
    bkw = dict(k=50, leaf_max=400, small_max=3000, iterations=1, rounds=2, non_passes=1, seed=9)
    Xb = X[:12000]
    ids0, ds0 = K.Builder(K.params(**bkw)).build(Xb, return_dists=True)
    ids1, ds1 = K.Builder(K.params(flags=K.FLAG_AUTOTUNE, **bkw)).build(Xb, return_dists=True)
    np.testing.assert_array_equal(ids1, ids0)
    np.testing.assert_array_equal(ds1, ds0)


@pytest.mark.parametrize("flags", [K.FLAG_LEAF_BATCHED, K.FLAG_LEAF_GRAPH])
def test_batched_leaf_stage_equals_binned_launches(flags):
    # One persistent kernel over the round's work list (and its graph replays) against one launch
    # per size bin: buckets up to 1500 points (blocks of the big-bucket path) and small ones, three
    # rounds in a row, each call's lists compared (a graph is replayed on every call after the
    # first, with host work and synchronisations between the replays).
    X = mixture(30000, 100, 60, 81)
    k = 50
    kw = dict(k=k, leaf_max=1500, small_max=3000, iterations=1, seed=8)
    o_prm = oracle.make_params(**kw)
    ref = K.Builder(K.params(**kw))
    bat = K.Builder(K.params(flags=flags, **kw))
    st_r = st_b = None
    for rnd in range(3):
        perm, ls, lz = oracle.split_round(X, o_prm, rnd)
        if rnd == 0:
            assert lz.max() > 512 and lz.min() < 256
        st_r = ref.leaf_stage(X, perm, ls, lz, state=st_r)
        st_b = bat.leaf_stage(X, perm, ls, lz, state=st_b)
        np.testing.assert_array_equal(st_b[1], st_r[1])
        np.testing.assert_array_equal(st_b[0], st_r[0])
    # whole builds: rounds alternate between two graphs (one per round parity)
    bkw = dict(k=50, leaf_max=1200, leaf_max_odd=300, small_max=3000, iterations=1, rounds=5,
               non_passes=1, seed=3)
    ids0, ds0 = K.Builder(K.params(**bkw)).build(X, return_dists=True)
    ids1, ds1 = K.Builder(K.params(flags=flags, **bkw)).build(X, return_dists=True)
    np.testing.assert_array_equal(ds1, ds0)
    np.testing.assert_array_equal(ids1, ids0)


@pytest.mark.parametrize("flags", [0, K.FLAG_DIRECT_NON])
def test_non_matches_oracle(flags):
    X = mixture(6000, 100, 30, 31)
    k = 20
    kw = dict(k=k, leaf_max=100, small_max=3000, iterations=1, seed=2)
    o_prm = oracle.make_params(**kw)
    perm, ls, lz = oracle.split_round(X, o_prm, 0)
    st = oracle.leaf_stage(X, k, perm, ls, lz)
    ids, ds = K.Builder(K.params(flags=flags, **kw)).non(X, (st[0], st[1]))
    o_ids, o_ds, _ = oracle.non(X, k, st)
    assert_lists_equal_up_to_ties(X, ids, ds, o_ids, o_ds)


@pytest.mark.parametrize("gen,d", [("lowrank", 64), ("mixture", 100), ("lowrank", 100), ("lowrank", 128),
                                   ("mixture", 37), ("lowrank_tiny", 100), ("mixture_huge", 64)])
@pytest.mark.parametrize("flags", [0, K.FLAG_NON_FP32_GRAM, K.FLAG_NON_SCREEN16])
def test_non_screens_match_oracle(gen, d, flags):
    # the reverse kernel's screens, each with its exact margin: the fp16 Gram (default for widths
    # 64, 100, 128) of the rows as they are when every squared norm lies in [2^-12, 2^28]
    # (unit-norm and mixture points), else of the rows scaled by powers of two (the tiny and huge
    # sets); the FP32 Gram (every width, the only one for others); the 16-bit Gram of rows centred
    # on each q. The lists must be the oracle's.
    if gen == "lowrank_tiny":
        X = lowrank(8000, d, 70 + d) * np.float32(2.0 ** -8)   # squared norms 2^-16
    elif gen == "mixture_huge":
        X = mixture(8000, d, 40, 70 + d) * np.float32(2.0 ** 12)  # squared norms ~2^34
    else:
        X = lowrank(8000, d, 70 + d) if gen == "lowrank" else mixture(8000, d, 40, 70 + d)
    k = 50
    kw = dict(k=k, leaf_max=128, small_max=3000, iterations=1, seed=12)
    o_prm = oracle.make_params(**kw)
    perm, ls, lz = oracle.split_round(X, o_prm, 0)
    st = oracle.leaf_stage(X, k, perm, ls, lz)
    ids, ds = K.Builder(K.params(flags=flags, **kw)).non(X, (st[0], st[1]))
    o_ids, o_ds, _ = oracle.non(X, k, st)
    assert_lists_equal_up_to_ties(X, ids, ds, o_ids, o_ds)


def test_non_overflow_path_matches_oracle():
    # 16 candidate slots per point: most buffers overflow and those points are redone by the
    # direct kernel; the lists must still equal the oracle's
    X = mixture(6000, 100, 30, 37)
    k = 20
    kw = dict(k=k, leaf_max=100, small_max=3000, iterations=1, seed=4)
    o_prm = oracle.make_params(**kw)
    perm, ls, lz = oracle.split_round(X, o_prm, 0)
    st = oracle.leaf_stage(X, k, perm, ls, lz)
    b = K.Builder(K.params(flags=K.FLAG_SMALL_NON_CAP, **kw))
    ids, ds = b.non(X, (st[0], st[1]))
    assert b.stats()["non_overflow"] > 0
    o_ids, o_ds, _ = oracle.non(X, k, st)
    assert_lists_equal_up_to_ties(X, ids, ds, o_ids, o_ds)


@pytest.mark.parametrize("flags,leaf", [(0, 100), (K.FLAG_DIRECT_NON, 100), (0, 15), (K.FLAG_DIRECT_NON, 15)])
def test_non_ordered_matches_oracle(flags, leaf):
    # topUpSingleOrdered: only neighbours at ranks >= 10 of 20 expanded. Leaf 100 fills every list;
    # leaf 15 leaves them all short (filled entries only are expanded, padding never)
    X = mixture(6000, 100, 30, 53)
    k = 20
    kw = dict(k=k, leaf_max=leaf, small_max=3000, iterations=1, seed=6)
    o_prm = oracle.make_params(**kw)
    perm, ls, lz = oracle.split_round(X, o_prm, 0)
    st = oracle.leaf_stage(X, k, perm, ls, lz)
    ids, ds = K.Builder(K.params(flags=flags, non_rank_lo=10, **kw)).non(X, (st[0], st[1]))
    o_ids, o_ds, _ = oracle.non(X, k, st, rank_lo=10)
    assert_lists_equal_up_to_ties(X, ids, ds, o_ids, o_ds)


def test_build_ordered_non_matches_oracle():
    X = lowrank(15000, 100, 57)
    kw = dict(k=100, leaf_max=256, small_max=3000, iterations=1, rounds=2, non_passes=2,
              non_rank_lo=50, seed=59)
    ids, ds = K.Builder(K.params(**kw)).build(X, return_dists=True)
    o_ids, o_ds, _, _ = oracle.build_graph(X, oracle.make_params(**kw))
    assert_lists_equal_up_to_ties(X, ids, ds, o_ids, o_ds)


def test_non_k100_lowrank():
    # k = 100 lists on close unit-norm points: many 2-hop candidates per point, buffer overflow path
    X = lowrank(8000, 100, 33)
    k = 100
    kw = dict(k=k, leaf_max=256, small_max=3000, iterations=1, seed=5)
    o_prm = oracle.make_params(**kw)
    perm, ls, lz = oracle.split_round(X, o_prm, 0)
    st = oracle.leaf_stage(X, k, perm, ls, lz)
    b = K.Builder(K.params(**kw))
    ids, ds = b.non(X, (st[0], st[1]))
    o_ids, o_ds, _ = oracle.non(X, k, st)
    assert_lists_equal_up_to_ties(X, ids, ds, o_ids, o_ds)


def test_non_with_unfilled_lists():
    # lists shorter than k: unfilled slots read as id 0 (KnnSets.hpp:29), as in the reference
    X = mixture(3000, 32, 10, 5)
    k = 40
    kw = dict(k=k, leaf_max=20, small_max=500, iterations=1, seed=3)
    o_prm = oracle.make_params(**kw)
    perm, ls, lz = oracle.split_round(X, o_prm, 0)
    st = oracle.leaf_stage(X, k, perm, ls, lz)
    assert (st[2] < k).all()
    ids, ds = K.Builder(K.params(**kw)).non(X, (st[0], st[1]))
    o_ids, o_ds, _ = oracle.non(X, k, st)
    assert_lists_equal_up_to_ties(X, ids, ds, o_ids, o_ds)


@pytest.mark.parametrize("k,rounds,non", [(20, 3, 1), (100, 2, 1), (10, 4, 0)])
def test_build_end_to_end_matches_oracle(k, rounds, non):
    X = mixture(15000, 100, 60, 41)
    kw = dict(k=k, leaf_max=256, small_max=3000, iterations=2, rounds=rounds, non_passes=non, seed=17)
    ids, ds = K.Builder(K.params(**kw)).build(X, return_dists=True)
    o_ids, o_ds, _, _ = oracle.build_graph(X, oracle.make_params(**kw))
    assert_lists_equal_up_to_ties(X, ids, ds, o_ids, o_ds)


@pytest.mark.parametrize("non", [0, 1])
def test_kd_build_matches_oracle(non):
    X = mixture(15000, 100, 60, 43)
    kw = dict(k=20, leaf_max=399, rounds=3, non_passes=non, seed=9)
    ids, ds = K.Builder(K.params(flags=K.FLAG_KD_SPLIT, **kw)).build(X, return_dists=True)
    o_ids, o_ds, _, _ = oracle.build_graph(X, oracle.make_params(splitter=1, **kw))
    assert_lists_equal_up_to_ties(X, ids, ds, o_ids, o_ds)


def test_build_reference_stride_and_padding():
    # stride-112 rows as the reference stores them (io.h:77), short lists padded with id 1
    X = mixture(2000, 100, 10, 6)
    Xs = np.zeros((2000, 112), np.float32)
    Xs[:, :100] = X
    b = K.Builder(K.params(k=30, leaf_max=10, small_max=200, rounds=1, non_passes=0, seed=1))
    ids = np.empty((2000, 30), np.uint32)
    ds = np.empty((2000, 30), np.float32)
    K._check(K.lib().knng_build(b.h, K._f(Xs), 2000, 100, 112, K._u(ids), K._f(ds)))
    o_ids, o_ds, o_sz, _ = oracle.build_graph(X, oracle.make_params(k=30, leaf_max=10, small_max=200,
                                                                     rounds=1, seed=1))
    assert_lists_equal_up_to_ties(X, ids, ds, o_ids, o_ds)
    short = o_sz < 30
    assert short.any()
    for r in np.where(short)[0]:
        assert np.all(ids[r, o_sz[r]:] == 1)


@pytest.mark.parametrize("d,stride", [(100, 112), (98, 100)])
def test_build_device_rows_ignore_padding(d, stride):
    # device rows used in place: dims d..stride-1 hold NaN and large values, which no kernel may
    # read into a distance, a screen or a split (the host path zero-pads its copy); with d = 98 the
    # padding shares the last 16-byte word of each row with dims 96 and 97
    n, k = 6000, 50
    X = mixture(n, d, 30, 8)
    Xs = np.empty((n, stride), np.float32)
    Xs[:, :d] = X
    Xs[:, d:] = np.float32(3e4)
    Xs[::3, d:] = np.nan
    kw = dict(k=k, leaf_max=256, small_max=2000, iterations=1, rounds=3, non_passes=2, seed=4)
    ids_h, ds_h = K.Builder(K.params(**kw)).build(X, return_dists=True)
    dX = K.DeviceArray(Xs.nbytes)
    dX.upload(Xs)
    dI = K.DeviceArray(n * k * 4)
    dD = K.DeviceArray(n * k * 4)
    K.Builder(K.params(**kw)).build_device(dX.ptr, n, d, stride, dI.ptr, dD.ptr)
    np.testing.assert_array_equal(dI.download((n, k), np.uint32), ids_h)
    np.testing.assert_array_equal(dD.download((n, k), np.float32), ds_h)


# ---------------------------------------------------------------------------------------------
# Precision forms, exact FP64 tool, generators
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("prec", [K.PREC_FP16, K.PREC_BF16])
@pytest.mark.parametrize("flags", [0, K.FLAG_NON_SCREEN16])
def test_precision_forms_valid_and_close(prec, flags):
    X = lowrank(20000, 100, 12)
    kw = dict(k=50, leaf_max=256, small_max=3000, iterations=1, rounds=4, non_passes=1, seed=3)
    ids, ds = K.Builder(K.params(precision=prec, flags=flags, **kw)).build(X, return_dists=True)
    f_ids, f_ds = K.Builder(K.params(**kw)).build(X, return_dists=True)
    assert_valid_lists(X, ids, ds)
    q = np.arange(0, 20000, 40, dtype=np.uint32)
    truth, _ = oracle.exact_knn(X, q, 50)
    r_p, r_f = K.recall(ids[q], truth), K.recall(f_ids[q], truth)
    assert r_p > r_f - 0.02, (r_p, r_f)


@pytest.mark.parametrize("k,leaf,d,flags", [(100, 400, 100, 0), (10, 400, 64, 0), (100, 1500, 100, 0),
                                             (100, 400, 128, K.FLAG_BIG_LEAF), (32, 2000, 37, 0)])
def test_leaf_bucket_limit_keeps_lists_and_cuts_survivors(k, leaf, d, flags):
    # The exact screen of the LDS leaf kernel lowers a row's bound to the k-th smallest upper
    # bound of its distances in the bucket (no pair of the k nearest is dropped): the lists must
    # be those of the screen without it, bit for bit, on empty lists (where it applies) and on a
    # second round merged into them, with fewer pairs re-ranked where every bucket is LDS-resident;
    # and both must match the oracle. Buckets above 512 points (or forced onto the big-bucket
    # kernel) take no limit in the exact screen: same lists, no more re-ranks.
    X = lowrank(16000, d, 31) if d != 37 else mixture(16000, d, 8, 31)
    kw = dict(k=k, leaf_max=leaf, small_max=3000, iterations=1, seed=6)
    o_prm = oracle.make_params(**kw)
    on = K.Builder(K.params(flags=flags, **kw))
    off = K.Builder(K.params(flags=flags | K.FLAG_NO_BUCKET_LIMIT, **kw))
    perm, ls, lz = oracle.split_round(X, o_prm, 0)
    ids1, ds1 = on.leaf_stage(X, perm, ls, lz)
    s_on = on.stats()["survivors"]
    ids0, ds0 = off.leaf_stage(X, perm, ls, lz)
    s_off = off.stats()["survivors"]
    np.testing.assert_array_equal(ds1, ds0)
    np.testing.assert_array_equal(ids1, ids0)
    assert s_on < 0.8 * s_off if leaf <= 512 and not flags else s_on <= s_off, (s_on, s_off)
    o_ids, o_ds, o_sz = oracle.leaf_stage(X, k, perm, ls, lz)
    assert_lists_equal_up_to_ties(X, ids1, ds1, o_ids, o_ds)
    perm2, ls2, lz2 = oracle.split_round(X, o_prm, 1)
    a1 = on.leaf_stage(X, perm2, ls2, lz2, state=(ids1, ds1))
    a0 = off.leaf_stage(X, perm2, ls2, lz2, state=(ids0, ds0))
    np.testing.assert_array_equal(a1[1], a0[1])
    np.testing.assert_array_equal(a1[0], a0[0])
    o2 = oracle.leaf_stage(X, k, perm2, ls2, lz2, state=(o_ids, o_ds, o_sz))
    assert_lists_equal_up_to_ties(X, a1[0], a1[1], o2[0], o2[1])


@pytest.mark.parametrize("prec", [K.PREC_FP16, K.PREC_BF16])
@pytest.mark.parametrize("flags", [0, K.FLAG_BIG_LEAF])
def test_rerank_cap_of_the_approximate_forms(prec, flags):
    # Approximate screens re-rank, per point and bucket, about the `rerank` nearest by the screen:
    # fewer pairs than every survivor (rerank = 0), every returned distance still the exact FP32
    # one, and recall against FP64 within 0.01 of the uncapped form's.
    X = lowrank(20000, 100, 14)
    kw = dict(k=100, leaf_max=400, small_max=3000, iterations=1, rounds=3, non_passes=0, seed=2)
    # one leaf stage on empty lists, where every pair passes the bound: about 160 re-ranks a point
    perm, ls, lz = oracle.split_round(X, oracle.make_params(**kw), 0)
    capped = K.Builder(K.params(precision=prec, flags=flags, rerank=160, **kw))
    full = K.Builder(K.params(precision=prec, flags=flags, rerank=0, **kw))
    capped.leaf_stage(X, perm, ls, lz)
    full.leaf_stage(X, perm, ls, lz)
    s_cap, s_full = capped.stats()["survivors"], full.stats()["survivors"]
    # every pair passes an empty list's bound; the big-bucket kernel tightens the bound between
    # its column blocks, so it re-ranks fewer
    all_pairs = int((lz.astype(np.int64) * (lz - 1)).sum())
    assert s_full == all_pairs if not flags else s_full <= all_pairs
    assert s_cap < 1.1 * 160 * 20000 and s_cap < 0.75 * s_full, (s_cap, s_full)
    ids, ds = capped.build(X, return_dists=True)
    f_ids, _ = full.build(X, return_dists=True)
    assert_valid_lists(X, ids, ds)
    q = np.arange(0, 20000, 40, dtype=np.uint32)
    truth, _ = oracle.exact_knn(X, q, 100)
    r_c, r_f = K.recall(ids[q], truth), K.recall(f_ids[q], truth)
    assert r_c > r_f - 0.01, (r_c, r_f)


def test_exact_knn_gpu_matches_oracle():
    X = mixture(4000, 100, 20, 51)
    q = np.arange(0, 4000, 37, dtype=np.uint32)
    b = K.Builder(K.params(k=10))
    ids, ds = b.exact_knn(X, q, 25)
    o_ids, o_ds = oracle.exact_knn(X, q, 25)
    np.testing.assert_allclose(ds, o_ds, rtol=1e-12, atol=1e-9)
    np.testing.assert_array_equal(ids, o_ids)


def test_recall_against_fp64_c0_like():
    X = mixture(20000, 100, 100, 61)
    kw = dict(k=100, leaf_max=256, small_max=3000, iterations=2, rounds=8, non_passes=1, seed=1)
    ids = K.Builder(K.params(**kw)).build(X)
    q = np.arange(0, 20000, 50, dtype=np.uint32)
    truth, _ = oracle.exact_knn(X, q, 100)
    assert K.recall(ids[q], truth) > 0.9


def test_generators_on_device():
    b = K.Builder(K.params(k=10))
    n, d = 4096, 100
    dX = K.DeviceArray(n * d * 4)
    b.generate_device(K.GEN_LOWRANK, n, d, d, 5, dX.ptr)
    X = dX.download((n, d), np.float32)
    np.testing.assert_allclose(np.linalg.norm(X, axis=1), 1.0, rtol=1e-5)
    # low rank: 24 latent directions carry almost all the energy
    s = np.linalg.svd(X - X.mean(0), compute_uv=False)
    assert (s[:24] ** 2).sum() / (s ** 2).sum() > 0.95
    b.generate_device(K.GEN_MIXTURE, n, d, d, 5, dX.ptr)
    Y = dX.download((n, d), np.float32)
    assert np.abs(Y).max() < 20 and np.abs(Y).max() > 8
    b.generate_device(K.GEN_LOWRANK, n, d, d, 5, dX.ptr)
    np.testing.assert_array_equal(dX.download((n, d), np.float32), X)     # seeded


def test_search_width_outputs_first_k_of_wider_lists():
    # k = 10 searched with 100-wide lists is the first 10 columns of the k = 100 build
    X = lowrank(20000, 64, 31)
    kw = dict(leaf_max=256, small_max=3000, iterations=1, rounds=3, non_passes=1, seed=29)
    ids10, ds10 = K.Builder(K.params(k=10, search_k=100, **kw)).build(X, return_dists=True)
    ids100, ds100 = K.Builder(K.params(k=100, **kw)).build(X, return_dists=True)
    np.testing.assert_array_equal(ids10, ids100[:, :10])
    np.testing.assert_array_equal(ds10, ds100[:, :10])


def test_incremental_non_passes_equal_full_passes():
    # later passes skip offers whose two hops both predate the previous pass: same lists exactly
    X = lowrank(30000, 100, 37)
    kw = dict(k=100, leaf_max=256, small_max=3000, iterations=1, rounds=2, non_passes=3, seed=23)
    ids_i, ds_i = K.Builder(K.params(flags=K.FLAG_INCREMENTAL_NON, **kw)).build(X, return_dists=True)
    ids_f, ds_f = K.Builder(K.params(**kw)).build(X, return_dists=True)
    np.testing.assert_array_equal(ds_i, ds_f)
    np.testing.assert_array_equal(ids_i, ids_f)


def test_build_two_non_passes_matches_oracle():
    X = mixture(12000, 64, 50, 47)
    kw = dict(k=40, leaf_max=128, small_max=3000, iterations=1, rounds=2, non_passes=2, seed=41)
    ids, ds = K.Builder(K.params(**kw)).build(X, return_dists=True)
    o_ids, o_ds, _, _ = oracle.build_graph(X, oracle.make_params(**kw))
    assert_lists_equal_up_to_ties(X, ids, ds, o_ids, o_ds)


@pytest.mark.timeout(600)
def test_full_size_lowrank_build_properties():
    # The headline workload (10M x 100 low-rank unit-norm points generated on the device, k = 100,
    # leaf <= 400, 6 rounds, 2 NoN passes), checked through properties that hold at any size:
    # rows sorted by (dist, id) with distinct ids, no self and no padding; sampled rows' distances
    # equal to the oracle's reference-order FP32 distance bit for bit; recall@100 against exact
    # FP64 k-NN (the device kernel that test_exact_knn_gpu_matches_oracle pins) above 0.94.
    n, d, k = 10_000_000, 100, 100
    b = K.Builder(K.params(k=k, leaf_max=400, small_max=3000, iterations=1, rounds=6,
                           non_passes=2, seed=2023))
    dX = K.DeviceArray(n * d * 4)
    b.generate_device(K.GEN_LOWRANK, n, d, d, 2023, dX.ptr)
    dI = K.DeviceArray(n * k * 4)
    dD = K.DeviceArray(n * k * 4)
    b.build_device(dX.ptr, n, d, d, dI.ptr, dD.ptr)
    ids = dI.download((n, k), np.uint32)
    ds = dD.download((n, k), np.float32)
    del dD
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
    X = dX.download((n, d), np.float32)
    rng = np.random.default_rng(3)
    for r in rng.choice(n, 200, replace=False):
        got = np.array([oracle.dist(X[r], X[j]) for j in ids[r]], np.float32)
        np.testing.assert_array_equal(ds[r], got)
    q = np.sort(rng.choice(n, 4000, replace=False)).astype(np.uint32)
    nq = q.size
    dq = K.DeviceArray(nq * 4)
    dq.upload(q)
    dl = K.DeviceArray(nq * k * 4)
    dl.upload(np.ascontiguousarray(ids[q]))
    dub = K.DeviceArray(nq * 8)
    b.exact_upper_from_lists(dX.ptr, n, d, d, dq.ptr, nq, dl.ptr, k, dub.ptr)
    dti = K.DeviceArray(nq * k * 4)
    dtd = K.DeviceArray(nq * k * 8)
    b.exact_knn_device(dX.ptr, n, d, d, dq.ptr, nq, k, dub.ptr, 4 * k + 256, dti.ptr, dtd.ptr)
    rec = K.recall(ids[q], dti.download((nq, k), np.uint32))
    assert rec > 0.94, rec
    del b, dX, dI
