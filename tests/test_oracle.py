"""Pins the CPU oracle (oracle/knng_oracle.cpp) before it is trusted as the parity checker.

Each check is chosen so that a plausible mistake (a dropped tail term, a wrong sign, a transposed
index, a reordered sum, a wrong tie rule) fails it: the reference's own known answers
(test/TestKnnSet.cpp), closed forms, exact integer arithmetic, targeted rounding cases, FP64 bounds,
and brute force on small inputs.
"""
import numpy as np
import pytest

import oracle

F32MAX = np.finfo(np.float32).max


# ----------------------------------------------------------------------------------------------
# distance / dot  (LinearAlgebra.hpp:53-77, :179-200)
# ----------------------------------------------------------------------------------------------
@pytest.mark.parametrize("d", [1, 7, 8, 64, 99, 100, 101, 128])
def test_dist_dot_integer_exact(d):
    # Integer inputs keep every partial sum exact: any dropped term, tail dim or sign error shows.
    rng = np.random.default_rng(d)
    a = rng.integers(-20, 21, d).astype(np.float32)
    b = rng.integers(-20, 21, d).astype(np.float32)
    assert oracle.dist(a, b) == np.float32(np.sum((a.astype(np.int64) - b.astype(np.int64)) ** 2))
    assert oracle.dot(a, b) == np.float32(np.sum(a.astype(np.int64) * b.astype(np.int64)))


def test_dist_tail_dims_counted():
    # Only dims 96..99 differ: the reference's scalar tail (LinearAlgebra.hpp:72-75).
    a = np.zeros(100, np.float32)
    b = np.zeros(100, np.float32)
    b[96:] = [1, 2, 3, 4]
    assert oracle.dist(a, b) == 30.0
    assert oracle.dot(b, b) == 30.0


@pytest.mark.parametrize("d", [100, 128, 64])
def test_dist_symmetric_zero_and_fp64_bound(d):
    rng = np.random.default_rng(7)
    for _ in range(50):
        a = rng.normal(size=d).astype(np.float32)
        b = rng.normal(size=d).astype(np.float32)
        dab = oracle.dist(a, b)
        assert dab == oracle.dist(b, a)          # bit-exact symmetry: (a-b)^2 == (b-a)^2
        assert oracle.dist(a, a) == 0.0
        exact = np.sum((a.astype(np.float64) - b.astype(np.float64)) ** 2)
        # fp32 recursive summation bound: |err| <= ~ (d/8 + 8 + 2) * u * sum
        assert abs(float(dab) - exact) <= (d / 8 + 12) * 2.0 ** -24 * exact + 1e-30
        dd = oracle.dot(a, b)
        exactd = np.dot(a.astype(np.float64), b.astype(np.float64))
        assert abs(float(dd) - exactd) <= (d / 8 + 12) * 2.0 ** -24 * np.dot(np.abs(a), np.abs(b))


def test_lane_sum_is_sequential():
    # Lane accumulators [2^24, 1, 1, 1, 1, 0, 0, 0]: summed 0..7 in order each +1 rounds away
    # (ties to even) and the result stays 2^24; any pairwise/tree order gives 2^24 + 4.
    a = np.array([2.0 ** 24, 1, 1, 1, 1, 0, 0, 0], np.float32)
    assert oracle.dot(a, np.ones(8, np.float32)) == np.float32(2.0 ** 24)


def test_lane_assignment_is_mod8():
    # Lane j accumulates dims j, j+8, ...: put 2^24 and two 1s into lane 0 (dims 0, 8, 16) so they
    # round away inside one lane accumulator; a different lane layout keeps them.
    a = np.zeros(24, np.float32)
    a[0] = 2.0 ** 24
    a[8] = 1.0
    a[16] = 1.0
    # lane 0: fma(1,1,fma(1,1,2^24)) = 2^24 (each +1 is a tie to even), other lanes 0
    assert oracle.dot(a, np.ones(24, np.float32)) == np.float32(2.0 ** 24)
    b = np.zeros(24, np.float32)
    b[0] = 2.0 ** 24
    b[1] = 1.0
    b[2] = 1.0   # different lanes: sum 0 + 2^24 + 1 + 1 sequential -> 2^24
    assert oracle.dot(b, np.ones(24, np.float32)) == np.float32(2.0 ** 24)
    c = np.zeros(24, np.float32)
    c[1] = 1.0
    c[2] = 1.0
    c[3] = 2.0 ** 24   # lanes 1, 2 first: 0 + 1 + 1 + 2^24 = 2^24 + 2 exactly
    assert oracle.dot(c, np.ones(24, np.float32)) == np.float32(2.0 ** 24 + 2)


def test_tail_is_fused():
    # dot tail ans = fma(a, b, ans): lanes give -1, tail (1+2^-12)^2 = 1 + 2^-11 + 2^-24.
    # fused: 2^-11 + 2^-24 ; separate multiply then add: 2^-11.
    a = np.zeros(9, np.float32)
    b = np.zeros(9, np.float32)
    a[0], b[0] = -1.0, 1.0
    t = np.float32(1.0 + 2.0 ** -12)
    a[8], b[8] = t, t
    assert oracle.dot(a, b) == np.float32(2.0 ** -11 + 2.0 ** -24)


# ----------------------------------------------------------------------------------------------
# Hyperplane algebra  (SolutionKmeans.hpp:162-164; LinearAlgebra.hpp:128-161)
# ----------------------------------------------------------------------------------------------
def test_hyperplane_closed_form():
    c1 = np.array([4, 2, 0, -2, 8, 6, 1, 3, 5], np.float32)
    c2 = np.array([0, 2, 4, 2, -8, 2, 1, 1, 3], np.float32)
    coef, off = oracle.hyperplane(c1, c2)
    np.testing.assert_array_equal(coef, (c1 - c2) / 2)          # exact for these integers
    mid = (c1 + c2) / 2
    assert off == np.float32(np.dot(mid.astype(np.float64), coef.astype(np.float64)))
    # c1 is on the lo side (dot >= offset), c2 on the hi side
    assert oracle.dot(coef, c1) >= off
    assert oracle.dot(coef, c2) < off


# ----------------------------------------------------------------------------------------------
# Sample sizes  (SolutionKmeans.hpp:50-58, :70, :105)
# ----------------------------------------------------------------------------------------------
@pytest.mark.parametrize("n", [3000, 10_000, 100_000, 2_000_000, 10_000_000])
def test_sample_count(n):
    perc = np.float32(np.power(np.log(float(n)) / np.log(30.0), 7) / n)
    assert oracle.sample_count(n) == int(np.float32(n) * perc)


def test_sample_count_known_values():
    # The reference comments: "0.005 for 1e7, around 10% for 10k" (SolutionKmeans.hpp:52) and
    # "129 samples for 10m bucket" (:70).
    assert abs(oracle.sample_count(10_000_000) / 1e7 - 0.005) < 0.001
    assert abs(oracle.sample_count(10_000) / 1e4 - 0.10) < 0.03
    assert oracle.start_count(10_000_000) == 129
    assert oracle.start_count(100_000) == 55


# ----------------------------------------------------------------------------------------------
# KnnSetScannableSimd known answers (test/TestKnnSet.cpp, restated for the hot-path type)
# ----------------------------------------------------------------------------------------------
def test_knnset_contains_not_full():                       # TestKnnSet.cpp:8-13
    ks = oracle.KnnSet(100)
    ks.offer(99, 0.1)
    assert ks.contains(99)
    assert not ks.contains(100)


def test_knnset_contains_full():                           # TestKnnSet.cpp:15-23
    ks = oracle.KnnSet(100)
    for i in range(100):
        ks.offer(i, np.float32(0.01 * i))
    assert ks.contains(50)
    assert not ks.contains(100)


def test_knnset_bound_while_filling():
    # KnnSets.hpp:40-47: the hot-path set publishes its bound only when it becomes full (the
    # legacy KnnSetScannable of TestKnnSet.cpp:25-32 tracked the max on every add).
    ks = oracle.KnnSet(100)
    for i in range(99):
        ks.offer(i, np.float32(0.01 * i))
        assert ks.bound == F32MAX
    ks.offer(99, np.float32(0.99))
    assert ks.bound == np.float32(0.99)


@pytest.mark.parametrize("seed", [123, 1, 2])
def test_knnset_bound_after_100(seed):                     # TestKnnSet.cpp:34-50
    data = [(np.float32(0.01 * i), i) for i in range(100)]
    rng = np.random.default_rng(seed)
    order = rng.permutation(100)
    ks = oracle.KnnSet(100)
    for j in order:
        ks.offer(data[j][1], data[j][0])
    ks.offer(100, np.float32(0.0001))
    assert ks.bound == np.float32(0.01 * 98)


@pytest.mark.parametrize("seed", [123, 5])
def test_knnset_finalize_1000(seed):                        # TestKnnSet.cpp:52-71
    rng = np.random.default_rng(seed)
    order = rng.permutation(1000)
    ks = oracle.KnnSet(100)
    for i in order:
        ks.offer(int(i), np.float32(0.01 * i))
    ids, d = ks.finalize()
    np.testing.assert_array_equal(ids, np.arange(100))


def test_knnset_dedupe_and_tie_rule():
    ks = oracle.KnnSet(8)
    for i in range(8):
        ks.offer(i, 1.0)
    assert ks.bound == 1.0
    assert not ks.offer(3, 0.5)                # already present: rejected (containsFull)
    # getMaxIdx on all-equal values keeps lane 0 / block 0: slot 0 is evicted first
    assert ks.max_idx() == 0
    assert ks.offer(100, 0.5)
    assert not ks.contains(0) and ks.contains(100)


def test_get_max_idx_lane_rule():
    # KnnSets.hpp:82-111: lane-wise strict > (earlier block wins inside a lane), then cross-lane
    # strict > from lane 0 (lower lane wins).
    ks = oracle.KnnSet(16)
    for s in range(16):
        ks.set_slot(s, s, 0.5)
    ks.set_slot(10, 10, 2.0)    # lane 2, block 1
    ks.set_slot(3, 3, 2.0)      # lane 3, block 0
    assert ks.max_idx() == 10
    ks.set_slot(9, 9, 3.0)      # lane 1, block 1
    ks.set_slot(1, 1, 3.0)      # lane 1, block 0: same lane, earlier block wins
    assert ks.max_idx() == 1


def test_contains_vector():                                 # TestKnnSet.cpp:148-159
    assert oracle.contains(np.arange(1, 17), 5)
    assert oracle.contains(np.arange(1, 12), 10)
    assert not oracle.contains(np.arange(1, 12), 123)


def test_knnset_matches_sort_semantics():
    # k smallest distinct ids among all offers (ties impossible here: distinct distances).
    rng = np.random.default_rng(3)
    offers = [(int(rng.integers(0, 400)), np.float32(rng.random())) for _ in range(2000)]
    dist_of = {}
    for i, d in offers:
        dist_of.setdefault(i, d)
    offers = [(i, dist_of[i]) for i, _ in offers]     # same id -> same distance, as in a graph
    ks = oracle.KnnSet(50)
    for i, d in offers:
        ks.offer(i, d)
    ids, _ = ks.finalize()
    expect = sorted(dist_of.items(), key=lambda t: (t[1], t[0]))[:50]
    np.testing.assert_array_equal(ids, [i for i, _ in expect])


# ----------------------------------------------------------------------------------------------
# Split  (SolutionKmeans.hpp:142-311)
# ----------------------------------------------------------------------------------------------
def _mixture(n, d, centres, seed, spread=10.0):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, size=(centres, d)).astype(np.float32)
    lab = rng.integers(0, centres, n)
    return (c[lab] + rng.normal(size=(n, d))).astype(np.float32)


def test_partition_injected_invariants():
    X = _mixture(3000, 100, 20, 1)
    rng = np.random.default_rng(2)
    perm0 = rng.permutation(3000).astype(np.uint32)
    starts = np.array([0, 1000, 1700], np.uint32)
    sizes = np.array([1000, 700, 1300], np.uint32)
    coefs = rng.normal(size=(3, 100)).astype(np.float32)
    offs = rng.normal(size=3).astype(np.float32)
    perm, lo = oracle.partition_injected(X, perm0, starts, sizes, coefs, offs)
    for s, z, c, o, nl in zip(starts, sizes, coefs, offs, lo):
        seg0 = perm0[s:s + z]
        side = np.array([oracle.dot(c, X[i]) >= o for i in seg0])
        expect = np.concatenate([seg0[side], seg0[~side]])    # stable: order kept in each side
        np.testing.assert_array_equal(perm[s:s + z], expect)
        assert nl == side.sum()


def _check_round(X, perm, ls, lz, leaf_max):
    n = X.shape[0]
    assert np.array_equal(np.sort(perm), np.arange(n))
    assert ls[0] == 0 and np.all(ls[1:] == ls[:-1] + lz[:-1]) and ls[-1] + lz[-1] == n
    assert lz.max() <= leaf_max and lz.min() >= 1


@pytest.mark.parametrize("small_max", [3000, 500])
def test_split_round_invariants_and_determinism(small_max):
    X = _mixture(20000, 100, 50, 3)
    p = oracle.make_params(leaf_max=256, small_max=small_max, iterations=2, seed=99)
    perm, ls, lz = oracle.split_round(X, p, 0, threads=8)
    _check_round(X, perm, ls, lz, 256)
    perm1, ls1, lz1 = oracle.split_round(X, p, 0, threads=1)
    np.testing.assert_array_equal(perm, perm1)
    np.testing.assert_array_equal(lz, lz1)
    perm2, _, _ = oracle.split_round(X, p, 1, threads=8)
    assert not np.array_equal(perm, perm2)     # a new round draws new hyperplanes


def test_split_groups_nearby_points():
    # Well separated clusters: almost every leaf holds points of one cluster.
    rng = np.random.default_rng(4)
    c = rng.uniform(-100, 100, size=(40, 32)).astype(np.float32)
    lab = rng.integers(0, 40, 8000)
    X = (c[lab] + rng.normal(size=(8000, 32))).astype(np.float32)
    p = oracle.make_params(leaf_max=400, iterations=1, seed=5)
    perm, ls, lz = oracle.split_round(X, p, 0)
    # a leaf of <= 400 may hold two 200-point clusters; it should not mix many
    mixed = sum(len(set(lab[perm[s:s + z]])) > 2 for s, z in zip(ls, lz))
    assert mixed <= 0.1 * len(ls)


def test_split_degenerate_identical_points():
    # All points equal: every attempt fails (farthest pair has distance 0); the fallback halves.
    X = np.ones((1000, 16), np.float32)
    p = oracle.make_params(leaf_max=100, small_max=300, seed=1)
    perm, ls, lz = oracle.split_round(X, p, 0)
    _check_round(X, perm, ls, lz, 100)


# ----------------------------------------------------------------------------------------------
# Random-rotation KD splitter (SolutionRandomKD.hpp:258-428)
# ----------------------------------------------------------------------------------------------
_M64 = (1 << 64) - 1


def _mix64(z):
    z &= _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


# ----------------------------------------------------------------------------------------------
# 2-means split decisions: start pair, Lloyd step, sample-only Lloyd
# (SolutionKmeans.hpp:68-138, :161-186, :218-260)
# ----------------------------------------------------------------------------------------------
def test_farthest_pair_first_maximum_wins():
    # Unit square corners: the diagonals (0,3) and (1,2) tie at distance 2; the first in (i, j)
    # loop order wins (strict >, SolutionKmeans.hpp:88-93).
    X = np.array([[0, 0], [1, 0], [0, 1], [1, 1]], np.float32)
    assert oracle.farthest_pair(X, [0, 1, 2, 3]) == (0, 3)
    assert oracle.farthest_pair(X, [3, 2, 1, 0]) == (0, 3)
    assert oracle.farthest_pair(X, [1, 0, 3, 2]) == (0, 3)
    # a repeated row: the pair of two copies never wins over a real distance
    assert oracle.farthest_pair(X, [0, 0, 1]) == (0, 2)
    # every distance 0: no pair beyond FLT_MIN (the reference would dereference null)
    assert oracle.farthest_pair(X, [2, 2, 2, 2]) is None
    # a subnormal distance (1e-40 < FLT_MIN) does not count either: the search starts at FLT_MIN
    Y = np.array([[0.0, 0.0], [1e-20, 0.0]], np.float32)
    assert 0.0 < oracle.dist(Y[0], Y[1]) < np.finfo(np.float32).tiny
    assert oracle.farthest_pair(Y, [0, 1]) is None
    # brute force on random rows, with a duplicated maximal pair planted later in the order
    rng = np.random.default_rng(3)
    Z = rng.normal(size=(40, 9)).astype(np.float32)
    rows = list(rng.permutation(40)[:30])
    best, arg = np.float32(np.finfo(np.float32).tiny), None
    for i in range(len(rows)):
        for j in range(i + 1, len(rows)):
            dd = oracle.dist(Z[rows[i]], Z[rows[j]])
            if dd > best:
                best, arg = dd, (i, j)
    assert oracle.farthest_pair(Z, rows) == arg
    i, j = arg
    assert oracle.farthest_pair(Z, rows + [rows[i], rows[j]]) == arg


_PHI = 0x9E3779B97F4A7C15


def _draw_below(akey, i, n):
    h = _mix64(akey + _PHI * (i + 1))
    return ((h >> 32) * n) >> 32


def _attempt_key(seed, rnd, attempt):
    key = _mix64(seed + _PHI * (rnd + 1))
    return _mix64(key ^ ((0xD1B54A32D192ED03 * (attempt + 1)) & _M64))


def _first_max_pair(X, rows):
    best, arg = np.float32(np.finfo(np.float32).tiny), None
    for i in range(len(rows)):
        for j in range(i + 1, len(rows)):
            dd = oracle.dist(X[rows[i]], X[rows[j]])
            if dd > best:
                best, arg = dd, (i, j)
    return arg


def _two_integer_clusters(rng, per, d, spread):
    """Two clusters of 2*per points with integer centres a, b; offsets come in +e/-e pairs, so
    each cluster's mean is its centre exactly. Points are interleaved: even rows in a."""
    a = rng.integers(-10, 11, d).astype(np.float64)
    b = a + rng.choice([-1, 1], d) * rng.integers(40, 60, d)
    e = rng.integers(1, spread + 1, (per, d)) * rng.choice([-1, 1], (per, d))
    A = np.concatenate([a + e, a - e])
    B = np.concatenate([b + e[::-1], b - e[::-1]])
    X = np.empty((4 * per, d), np.float32)
    X[0::2] = A
    X[1::2] = B
    return X, a, b


@pytest.mark.parametrize("iterations", [1, 2])
def test_split_root_lloyd_step_closed_form(iterations):
    # Small branch (n < small_max), full-range Lloyd with FP64 sums (SolutionKmeans.hpp:161-186).
    # After one step the centres are the two cluster means exactly, so the hyperplane is
    # coef = (m1 - m2) / 2, offset = dot((m1 + m2) / 2, (m1 - m2) / 2), with m1 the mean of the
    # cluster holding the first start point; the lo child is that cluster, in index order.
    d, seed = 24, 17
    X, a, b = _two_integer_clusters(np.random.default_rng(11), 32, d, 3)
    n = X.shape[0]
    p = oracle.make_params(leaf_max=32, small_max=3000, iterations=iterations, seed=seed)
    perm, t = oracle.split_root_trace(X, p, 0)
    # the start sample is the first s draws of attempt 0 (SolutionKmeans.hpp:70-74, :103-110)
    assert t["attempt"] == 0 and not t["fallback"]
    s = oracle.start_count(n)
    rows = [_draw_below(_attempt_key(seed, 0, 0), i, n) for i in range(s)]
    np.testing.assert_array_equal(t["rows"], rows)
    bi, bj = _first_max_pair(X, rows)
    assert (t["bi"], t["bj"]) == (bi, bj)
    ri, rj = rows[bi], rows[bj]
    assert ri % 2 != rj % 2, "precondition: the start pair straddles the clusters"
    m1, m2 = (a, b) if ri % 2 == 0 else (b, a)
    np.testing.assert_array_equal(t["c1"], m1.astype(np.float32))
    np.testing.assert_array_equal(t["c2"], m2.astype(np.float32))
    np.testing.assert_array_equal(t["coef"], ((m1 - m2) / 2).astype(np.float32))
    assert t["offset"] == np.float32(np.dot((m1 + m2) / 2, (m1 - m2) / 2))   # exact: quarter-integers
    lo = np.arange(ri % 2, n, 2)
    hi = np.arange(1 - ri % 2, n, 2)
    assert t["nlo"] == lo.size
    np.testing.assert_array_equal(perm, np.concatenate([lo, hi]))


def test_split_root_start_pair_hyperplane_without_lloyd():
    # iterations = 0: the hyperplane is the bisector of the farthest start pair itself.
    rng = np.random.default_rng(5)
    X = rng.integers(-8, 9, (700, 12)).astype(np.float32)
    seed = 3
    p = oracle.make_params(leaf_max=64, small_max=3000, iterations=0, seed=seed)
    perm, t = oracle.split_root_trace(X, p, 2)
    n = X.shape[0]
    rows = [_draw_below(_attempt_key(seed, 2, t["attempt"]), i, n) for i in range(oracle.start_count(n))]
    np.testing.assert_array_equal(t["rows"], rows)
    bi, bj = _first_max_pair(X, rows)
    assert (t["bi"], t["bj"]) == (bi, bj)
    xa, xb = X[rows[bi]].astype(np.float64), X[rows[bj]].astype(np.float64)
    np.testing.assert_array_equal(t["coef"], ((xa - xb) / 2).astype(np.float32))
    assert t["offset"] == np.float32(np.dot((xa + xb) / 2, (xa - xb) / 2))


def test_split_root_large_branch_lloyd_uses_the_sample_only():
    # Large branch (n >= small_max): the Lloyd step sums only the m sample draws, with
    # replacement (SolutionKmeans.hpp:218-249); the final side test covers the whole range
    # (:262-289). Integer coordinates keep every FP64 sum and every start-plane dot exact, so the
    # expected centres are computed here independently of the summation order.
    rng = np.random.default_rng(21)
    n, d, seed = 5000, 16, 8
    X = rng.integers(-4, 5, (n, d)).astype(np.float32)
    p = oracle.make_params(leaf_max=400, small_max=3000, iterations=1, seed=seed)
    perm, t = oracle.split_root_trace(X, p, 0)
    assert t["attempt"] == 0 and not t["fallback"]
    m = oracle.sample_count(n)
    assert t["m"] == m and m < n
    ak = _attempt_key(seed, 0, 0)
    draws = np.array([_draw_below(ak, i, n) for i in range(m)])
    s = min(oracle.start_count(n), m)
    np.testing.assert_array_equal(t["rows"], draws[:s])
    bi, bj = _first_max_pair(X, list(draws[:s]))
    assert (t["bi"], t["bj"]) == (bi, bj)
    xa, xb = X[draws[bi]].astype(np.float64), X[draws[bj]].astype(np.float64)
    coef0, off0 = (xa - xb) / 2, np.dot((xa + xb) / 2, (xa - xb) / 2)

    def means(idx):
        Y = X[idx].astype(np.float64)
        side = Y @ coef0 >= off0
        return ((Y[side].sum(0) / side.sum()).astype(np.float32),
                (Y[~side].sum(0) / (~side).sum()).astype(np.float32))

    c1, c2 = means(draws)
    f1, f2 = means(np.arange(n))
    assert not np.array_equal(c1, f1), "precondition: sample and full-range means differ"
    np.testing.assert_array_equal(t["c1"], c1)
    np.testing.assert_array_equal(t["c2"], c2)
    between = np.float32(0.5) * (c1 + c2)
    coef = c1 - between
    np.testing.assert_array_equal(t["coef"], coef)
    assert t["offset"] == oracle.dot(between, coef)
    side = np.array([oracle.dot(coef, x) >= t["offset"] for x in X])
    idx = np.arange(n)
    np.testing.assert_array_equal(perm, np.concatenate([idx[side], idx[~side]]))
    assert t["nlo"] == side.sum()


def _kd_key(seed, rnd):
    rk = _mix64(seed + 0x9E3779B97F4A7C15 * (rnd + 1))
    return _mix64(rk ^ 0x6A09E667F3BCC909)


def test_kd_normal_is_box_muller_of_the_counter_stream():
    key = _kd_key(11, 2)
    for i in (0, 1, 7, 12345):
        a = _mix64(key + 0x9E3779B97F4A7C15 * (2 * i + 1))
        b = _mix64(key + 0x9E3779B97F4A7C15 * (2 * i + 2))
        u1 = ((a >> 11) + 1.0) * 2.0 ** -53
        u2 = (b >> 11) * 2.0 ** -53
        expect = np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)
        assert oracle.kd_normal(key, i) == pytest.approx(expect, rel=1e-14, abs=1e-14)
    z = np.array([oracle.kd_normal(key, i) for i in range(20000)])
    assert abs(z.mean()) < 0.03 and abs(z.var() - 1.0) < 0.04


@pytest.mark.parametrize("m,d", [(30, 100), (8, 37)])
def test_kd_directions_follow_the_reference_formula(m, d):
    # randUniformUnitVec then gramSchmidt with the reference's identity `normalize`
    # (LinearAlgebra.hpp:94-101): u_0 = v_0, u_i = v_i - sum_j (u_j . v_i) u_j, recomputed in FP64
    seed, rnd = 7, 3
    key = _kd_key(seed, rnd)
    v = np.array([[np.float32(oracle.kd_normal(key, c * d + j)) for j in range(d)] for c in range(m)],
                 np.float64)
    v = v / np.sqrt((v * v).sum(axis=1, keepdims=True))
    u = np.zeros_like(v)
    for i in range(m):
        u[i] = v[i] - sum(np.dot(u[j], v[i]) * u[j] for j in range(i))
    got = oracle.kd_directions(seed, rnd, m, d).astype(np.float64)
    np.testing.assert_allclose(got, u, rtol=0, atol=2e-5)
    # not renormalised: the later rows are shorter than 1 (what the reference's bug produces)
    assert np.linalg.norm(got[0]) == pytest.approx(1.0, abs=1e-6)
    assert np.linalg.norm(got[-1]) < 0.999


def _pick_dim_py(var, key):
    m = len(var)
    tot = np.float32(0)
    for j in range(m):
        tot = np.float32(np.float64(tot) + var[j])
    order = sorted(((var[j], j) for j in range(m)), reverse=True)
    u01 = np.float32(_mix64(key + 0x9E3779B97F4A7C15) >> 40) * np.float32(2.0 ** -24)
    sample = np.float32(u01 * tot)
    so_far = np.float32(0)
    for vj, j in order:
        so_far = np.float32(np.float64(so_far) + vj)
        if so_far >= sample:
            return j
    return order[-1][1]


def test_kd_pick_dim_rule():
    rng = np.random.default_rng(8)
    hits = np.zeros(4)
    for t in range(3000):
        var = rng.uniform(0, 1, 4)
        if t % 7 == 0:
            var[1] = var[2]                     # ties: the larger index comes first
        key = int(rng.integers(0, 2 ** 63))
        j = oracle.kd_pick_dim(var, key)
        assert j == _pick_dim_py(var, key)
        hits[j] += var[j] / var.sum()
    assert oracle.kd_pick_dim(np.zeros(5), 3) == 4  # all zero: the draw is 0, the first in order
    # dimensions are drawn in proportion to their variance
    var = np.array([1.0, 3.0, 0.0, 6.0])
    cnt = np.bincount([oracle.kd_pick_dim(var, k) for k in range(20000)], minlength=4) / 20000
    np.testing.assert_allclose(cnt, var / var.sum(), atol=0.015)


def test_kd_split_round_root_is_a_mean_threshold():
    X = _mixture(6000, 100, 20, 12)
    p = oracle.make_params(leaf_max=399, splitter=1, seed=21)
    perm, ls, lz = oracle.split_round(X, p, 0)
    _check_round(X, perm, ls, lz, 399)
    perm1, _, lz1 = oracle.split_round(X, p, 0, threads=1)
    np.testing.assert_array_equal(perm, perm1)
    # the root's first child is exactly {x : rot_j(x) < mean_j} for one projected dimension j
    u = oracle.kd_directions(21, 0, 30, 100).astype(np.float64)
    rot = X.astype(np.float64) @ u.T
    # recover the root's lo size: the leaves that lie in the first child are those whose points all
    # fall below the threshold; search j and the split position together
    found = False
    for j in range(30):
        thr = rot[:, j].mean()
        lo = np.sort(np.where(rot[:, j] < thr)[0])
        nlo = len(lo)
        if 0 < nlo < 6000 and np.array_equal(np.sort(perm[:nlo]), lo):
            found = True
            break
    assert found
    perm2, _, _ = oracle.split_round(X, p, 1)
    assert not np.array_equal(perm, perm2)


def test_kd_split_identical_points_fall_back_to_halves():
    X = np.ones((2000, 16), np.float32)
    p = oracle.make_params(leaf_max=99, splitter=1, seed=2)
    perm, ls, lz = oracle.split_round(X, p, 0)
    _check_round(X, perm, ls, lz, 99)
    np.testing.assert_array_equal(perm, np.arange(2000))


# ----------------------------------------------------------------------------------------------
# Leaf stage (KnnSets.hpp:131-160) and NoN (SolutionKmeans.hpp:480-555) against brute force
# ----------------------------------------------------------------------------------------------
def _brute_sets(X, groups, k, old=None):
    """k smallest distinct (dist, id) per point among its group mates (plus its old set)."""
    n = X.shape[0]
    out = [dict() for _ in range(n)]
    if old is not None:
        ids, ds, sz = old
        for p in range(n):
            for s in range(sz[p]):
                out[p][int(ids[p, s])] = ds[p, s]
    for g in groups:
        for i in g:
            for j in g:
                if i != j and int(j) not in out[i]:
                    out[i][int(j)] = oracle.dist(X[i], X[j])
    res = []
    for p in range(n):
        items = sorted(out[p].items(), key=lambda t: (t[1], t[0]))[:k]
        res.append(items)
    return res


def _assert_state_equals(ids, ds, sz, expect):
    for p, items in enumerate(expect):
        assert sz[p] == len(items)
        np.testing.assert_array_equal(ds[p, :sz[p]], [d for _, d in items])
        np.testing.assert_array_equal(ids[p, :sz[p]], [i for i, _ in items])


def test_leaf_stage_brute_force():
    X = _mixture(1500, 24, 10, 8)
    k = 20
    p = oracle.make_params(k=k, leaf_max=60, small_max=400, seed=11)
    perm, ls, lz = oracle.split_round(X, p, 0)
    ids, ds, sz = oracle.leaf_stage(X, k, perm, ls, lz)
    groups = [perm[s:s + z] for s, z in zip(ls, lz)]
    _assert_state_equals(ids, ds, sz, _brute_sets(X, groups, k))
    # second round merges into the existing sets
    perm2, ls2, lz2 = oracle.split_round(X, p, 1)
    st2 = oracle.leaf_stage(X, k, perm2, ls2, lz2, state=(ids, ds, sz))
    groups2 = [perm2[s:s + z] for s, z in zip(ls2, lz2)]
    _assert_state_equals(*st2, _brute_sets(X, groups2, k, old=(ids, ds, sz)))


def test_non_brute_force():
    X = _mixture(600, 16, 6, 9)
    k = 10
    p = oracle.make_params(k=k, leaf_max=30, small_max=200, seed=3)
    perm, ls, lz = oracle.split_round(X, p, 0)
    st = oracle.leaf_stage(X, k, perm, ls, lz)
    ids, ds, sz = st
    new = oracle.non(X, k, st)
    # brute force: candidates = union of neighbours' snapshot lists (unfilled slots read as id 0)
    snap = np.where(np.arange(k)[None, :] < sz[:, None], ids, 0)
    expect = []
    for q in range(X.shape[0]):
        cur = {int(ids[q, s]): ds[q, s] for s in range(sz[q])}
        cands = set()
        for nb in snap[q]:
            cands.update(int(v) for v in snap[nb])
        cands.discard(q)
        for c in cands:
            if c not in cur:
                cur[c] = oracle.dist(X[c], X[q])
        expect.append(sorted(cur.items(), key=lambda t: (t[1], t[0]))[:k])
    _assert_state_equals(*new, expect)


@pytest.mark.parametrize("rank_lo", [5, 8])
def test_non_ordered_brute_force(rank_lo):
    # topUpSingleOrdered: only neighbours ranked >= rank_lo by (dist, id) are expanded, each into
    # its filled neighbours (first 100); self removed; result = k smallest of old + candidates
    X = _mixture(600, 16, 6, 19)
    k = 10
    p = oracle.make_params(k=k, leaf_max=30, small_max=200, seed=5)
    perm, ls, lz = oracle.split_round(X, p, 0)
    st = oracle.leaf_stage(X, k, perm, ls, lz)
    ids, ds, sz = st
    new = oracle.non(X, k, st, rank_lo=rank_lo)
    order = [[int(ids[q, s]) for s in sorted(range(sz[q]), key=lambda s: (ds[q, s], ids[q, s]))]
             for q in range(X.shape[0])]
    expect = []
    for q in range(X.shape[0]):
        cur = {int(ids[q, s]): ds[q, s] for s in range(sz[q])}
        cands = set()
        for nb in order[q][rank_lo:]:
            cands.update(order[nb][:100])
        cands.discard(q)
        for c in cands:
            if c not in cur:
                cur[c] = oracle.dist(X[c], X[q])
        expect.append(sorted(cur.items(), key=lambda t: (t[1], t[0]))[:k])
    _assert_state_equals(*new, expect)
    # expanding from rank 0 is a different pass from topUpSingle only through unfilled slots
    assert (sz == k).any()


# ----------------------------------------------------------------------------------------------
# Exact FP64 oracle and the whole build
# ----------------------------------------------------------------------------------------------
def test_exact_knn_numpy():
    X = _mixture(400, 20, 5, 10)
    q = np.array([0, 7, 399], np.uint32)
    ids, ds = oracle.exact_knn(X, q, 15)
    X64 = X.astype(np.float64)
    for r, qi in enumerate(q):
        dd = np.sum((X64 - X64[qi]) ** 2, axis=1)
        dd[qi] = np.inf
        order = np.lexsort((np.arange(400), dd))[:15]
        np.testing.assert_array_equal(ids[r], order)
        np.testing.assert_allclose(ds[r], dd[order], rtol=1e-12)


def test_build_recall_and_padding():
    X = _mixture(4000, 100, 40, 12)
    p = oracle.make_params(k=20, leaf_max=256, iterations=2, rounds=6, non_passes=1, seed=21)
    ids, ds, sz, st = oracle.build_graph(X, p)
    q = np.arange(0, 4000, 10, dtype=np.uint32)
    truth, _ = oracle.exact_knn(X, q, 20)
    assert oracle.recall(ids[q], truth) > 0.95
    assert np.all(ids[np.arange(4000), 0] != np.arange(4000))     # never self
    # padResult (KnnSets.hpp:334-344): short lists are padded with id 1
    p1 = oracle.make_params(k=20, leaf_max=8, small_max=50, rounds=1, seed=2)
    ids1, _, sz1, _ = oracle.build_graph(X[:200], p1)
    short = np.where(sz1 < 20)[0]
    assert short.size > 0
    for r in short:
        assert np.all(ids1[r, sz1[r]:] == 1)
