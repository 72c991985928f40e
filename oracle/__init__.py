"""ctypes view of the CPU restatement in ``oracle/knng_oracle.cpp``.

Test infrastructure only: tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of
``bench.py`` load it as the checker. Nothing in ``sigmod23ann_amd`` imports it.

Each wrapper names the reference lines its C function restates (paths relative to the
reference root). Functions whose parity is only pinned by invariants say so in their test.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force: bool = True) -> None:
    """Compile liboracle.so with the committed Makefile."""
    args = ["make", "-C", _HERE]
    if force:
        args.append("-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


class Params(C.Structure):
    _fields_ = [
        ("k", C.c_uint32),
        ("leaf_max", C.c_uint32),
        ("leaf_max_odd", C.c_uint32),
        ("small_max", C.c_uint32),
        ("iterations", C.c_uint32),
        ("rounds", C.c_uint32),
        ("non_passes", C.c_uint32),
        ("non_rank_lo", C.c_uint32),
        ("splitter", C.c_uint32),
        ("pad_", C.c_uint32),
        ("seed", C.c_uint64),
    ]


class SplitTrace(C.Structure):
    """One node's split decisions (the oracle's SplitTrace): the successful attempt, the start
    sample rows and farthest pair (indices into ``rows``), the Lloyd sample size ``m``, the final
    centres, the side-test hyperplane and the lo size."""
    _fields_ = [
        ("attempt", C.c_uint32),
        ("s", C.c_uint32),
        ("bi", C.c_uint32),
        ("bj", C.c_uint32),
        ("m", C.c_uint32),
        ("nlo", C.c_uint32),
        ("fallback", C.c_uint32),
        ("pad_", C.c_uint32),
        ("rows", C.c_uint32 * 256),
        ("c1", C.c_float * 512),
        ("c2", C.c_float * 512),
        ("coef", C.c_float * 512),
        ("offset", C.c_float),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("split_s", C.c_double),
        ("leaf_s", C.c_double),
        ("non_s", C.c_double),
        ("finalize_s", C.c_double),
        ("attempts_failed", C.c_uint64),
        ("fallbacks", C.c_uint64),
        ("leaves", C.c_uint64),
    ]


def make_params(k=100, leaf_max=256, leaf_max_odd=0, small_max=3000, iterations=1, rounds=1,
                non_passes=0, seed=1234, non_rank_lo=0, splitter=0, **_ignored) -> Params:
    """Reference build parameters; GPU-only knobs (flags, precision, search width) are ignored.
    splitter: 0 = 2-means (SolutionKmeans), 1 = random-rotation KD (SolutionRandomKD)."""
    return Params(k, leaf_max, leaf_max_odd, small_max, iterations, rounds, non_passes, non_rank_lo,
                  splitter, 0, seed)


_P = C.POINTER
_f32p = _P(C.c_float)
_u32p = _P(C.c_uint32)
_f64p = _P(C.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_dist.argtypes = [_f32p, _f32p, C.c_uint32]
        L.orc_dist.restype = C.c_float
        L.orc_dot.argtypes = [_f32p, _f32p, C.c_uint32]
        L.orc_dot.restype = C.c_float
        L.orc_sample_count.argtypes = [C.c_uint32]
        L.orc_sample_count.restype = C.c_uint32
        L.orc_start_count.argtypes = [C.c_uint32]
        L.orc_start_count.restype = C.c_uint32
        L.orc_mix64.argtypes = [C.c_uint64]
        L.orc_mix64.restype = C.c_uint64
        L.orc_draw_below.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32]
        L.orc_draw_below.restype = C.c_uint32
        L.orc_round_key.argtypes = [C.c_uint64, C.c_uint32]
        L.orc_round_key.restype = C.c_uint64
        L.orc_child_key.argtypes = [C.c_uint64, C.c_uint32]
        L.orc_child_key.restype = C.c_uint64
        L.orc_attempt_key.argtypes = [C.c_uint64, C.c_uint32]
        L.orc_attempt_key.restype = C.c_uint64
        L.orc_knnset_new.argtypes = [C.c_uint32]
        L.orc_knnset_new.restype = C.c_void_p
        L.orc_knnset_free.argtypes = [C.c_void_p]
        L.orc_knnset_offer.argtypes = [C.c_void_p, _f32p, C.c_uint32, C.c_float]
        L.orc_knnset_offer.restype = C.c_int
        L.orc_knnset_contains.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_knnset_contains.restype = C.c_int
        L.orc_knnset_size.argtypes = [C.c_void_p]
        L.orc_knnset_size.restype = C.c_uint32
        L.orc_knnset_max_idx.argtypes = [C.c_void_p]
        L.orc_knnset_max_idx.restype = C.c_uint32
        L.orc_knnset_set_slot.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_float]
        L.orc_knnset_finalize.argtypes = [C.c_void_p, _u32p, _f32p]
        L.orc_contains.argtypes = [_u32p, C.c_uint32, C.c_uint32]
        L.orc_contains.restype = C.c_int
        L.orc_split_round.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_uint32, _P(Params),
                                      C.c_uint32, _u32p, _u32p, _u32p, _u32p, C.c_int]
        L.orc_farthest_pair.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_uint32, _u32p,
                                        C.c_uint32, _u32p, _u32p]
        L.orc_farthest_pair.restype = C.c_int
        L.orc_split_root_trace.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_uint32, _P(Params),
                                           C.c_uint32, _u32p, _P(SplitTrace), C.c_int]
        L.orc_split_root_trace.restype = C.c_int
        L.orc_partition_injected.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_uint32, _u32p,
                                             C.c_uint32, _u32p, _u32p, _f32p, _f32p, _u32p]
        L.orc_hyperplane.argtypes = [_f32p, _f32p, C.c_uint32, _f32p]
        L.orc_hyperplane.restype = C.c_float
        L.orc_leaf_stage.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                     _u32p, _u32p, _u32p, C.c_uint32, _u32p, _f32p, _u32p,
                                     _u32p, _f32p, _u32p, C.c_int]
        L.orc_non.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _u32p,
                              _f32p, _u32p, _u32p, _f32p, _u32p, C.c_int, C.c_uint32]
        L.orc_build.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_uint32, _P(Params), _u32p,
                                _f32p, _u32p, C.c_int, _P(Stats)]
        L.orc_kd_directions.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, _f32p]
        L.orc_kd_normal.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_kd_normal.restype = C.c_double
        L.orc_kd_pick_dim.argtypes = [_f64p, C.c_uint32, C.c_uint64]
        L.orc_kd_pick_dim.restype = C.c_uint32
        L.orc_exact_knn.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_uint32, _u32p, C.c_uint32,
                                    C.c_uint32, _u32p, _f64p, C.c_int]
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(_f32p) if a is not None else None


def _up(a):
    return a.ctypes.data_as(_u32p) if a is not None else None


def _points(X):
    X = np.ascontiguousarray(X, dtype=np.float32)
    assert X.ndim == 2
    return X, X.shape[0], X.shape[1]


def dist(a, b) -> np.float32:
    """distance, LinearAlgebra.hpp:53-77."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return np.float32(lib().orc_dist(_fp(a), _fp(b), a.shape[0]))


def dot(a, b) -> np.float32:
    """dot, LinearAlgebra.hpp:179-200."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return np.float32(lib().orc_dot(_fp(a), _fp(b), a.shape[0]))


def hyperplane(c1, c2):
    """between/coefs/offset, SolutionKmeans.hpp:162-164."""
    c1 = np.ascontiguousarray(c1, np.float32)
    c2 = np.ascontiguousarray(c2, np.float32)
    coef = np.empty_like(c1)
    off = lib().orc_hyperplane(_fp(c1), _fp(c2), c1.shape[0], _fp(coef))
    return coef, np.float32(off)


def sample_count(n: int) -> int:
    """getSampleFromPercent size, SolutionKmeans.hpp:50-58."""
    return lib().orc_sample_count(n)


def start_count(n: int) -> int:
    """kmeansStartVecs sample size, SolutionKmeans.hpp:70, :105."""
    return lib().orc_start_count(n)


class KnnSet:
    """KnnSetScannableSimd (KnnSets.hpp:26-127) with its caller-side bound (``bounds[id]``)."""

    def __init__(self, k: int = 100):
        self.k = k
        self._h = lib().orc_knnset_new(k)
        self._bound = C.c_float(np.finfo(np.float32).max)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_knnset_free(self._h)
            self._h = None

    @property
    def bound(self) -> float:
        return float(self._bound.value)

    def offer(self, cand_id: int, dist: float) -> bool:
        return bool(lib().orc_knnset_offer(self._h, C.byref(self._bound), cand_id, dist))

    def contains(self, cand_id: int) -> bool:
        return bool(lib().orc_knnset_contains(self._h, cand_id))

    @property
    def size(self) -> int:
        return lib().orc_knnset_size(self._h)

    def max_idx(self) -> int:
        return lib().orc_knnset_max_idx(self._h)

    def set_slot(self, slot: int, cand_id: int, dist: float) -> None:
        lib().orc_knnset_set_slot(self._h, slot, cand_id, dist)

    def finalize(self):
        ids = np.zeros(self.size, np.uint32)
        d = np.zeros(self.size, np.float32)
        lib().orc_knnset_finalize(self._h, _up(ids), _fp(d))
        return ids, d


def contains(ids, cand: int) -> bool:
    """contains(vector, id), KnnSets.hpp:363-379."""
    ids = np.ascontiguousarray(ids, np.uint32)
    return bool(lib().orc_contains(_up(ids), ids.shape[0], cand))


def split_round(X, params: Params, round_idx: int, threads: int = 0):
    """One round of splitKmeansBinaryProcess (SolutionKmeans.hpp:142-311, :589-591).

    Returns (perm, leaf_start, leaf_size)."""
    X, n, d = _points(X)
    perm = np.empty(n, np.uint32)
    ls = np.empty(max(n, 1), np.uint32)
    lz = np.empty(max(n, 1), np.uint32)
    nl = C.c_uint32(0)
    lib().orc_split_round(_fp(X), n, d, d, C.byref(params), round_idx, _up(perm), _up(ls),
                          _up(lz), C.byref(nl), threads)
    return perm, ls[: nl.value].copy(), lz[: nl.value].copy()


def farthest_pair(X, rows):
    """kmeansStartVecs' pair search (SolutionKmeans.hpp:76-99) over X[rows]: the first pair
    (i, j), i < j, in loop order whose distance is the maximum and beyond FLT_MIN, as indices into
    ``rows``; None when there is none (the reference dereferences null there)."""
    X, n, d = _points(X)
    r = np.ascontiguousarray(rows, np.uint32)
    bi, bj = C.c_uint32(0), C.c_uint32(0)
    ok = lib().orc_farthest_pair(_fp(X), n, d, d, _up(r), r.shape[0], C.byref(bi), C.byref(bj))
    return (bi.value, bj.value) if ok else None


def split_root_trace(X, params: Params, round_idx: int = 0, threads: int = 0):
    """Splits the root node of a round once (SolutionKmeans.hpp:142-311) from the identity
    permutation. Returns (perm, trace dict): attempt, start rows, (bi, bj), m, c1, c2, coef,
    offset, nlo, fallback."""
    X, n, d = _points(X)
    perm = np.empty(n, np.uint32)
    tr = SplitTrace()
    rc = lib().orc_split_root_trace(_fp(X), n, d, d, C.byref(params), round_idx, _up(perm),
                                    C.byref(tr), threads)
    if rc != 0:
        raise ValueError("split_root_trace needs 2 <= n and d <= 512")
    t = dict(attempt=tr.attempt, s=tr.s, bi=tr.bi, bj=tr.bj, m=tr.m, nlo=tr.nlo,
             fallback=bool(tr.fallback), offset=np.float32(tr.offset),
             rows=np.array(tr.rows[: min(tr.s, 256)], np.uint32),
             c1=np.array(tr.c1[:d], np.float32), c2=np.array(tr.c2[:d], np.float32),
             coef=np.array(tr.coef[:d], np.float32))
    return perm, t


def kd_directions(seed: int, round_idx: int, m: int, d: int) -> np.ndarray:
    """The KD splitter's projection directions for a round (rotateRandom, SolutionRandomKD.hpp:258-264)."""
    out = np.empty((m, d), np.float32)
    lib().orc_kd_directions(seed, round_idx, m, d, _fp(out))
    return out


def kd_normal(key: int, i: int) -> float:
    return lib().orc_kd_normal(key, i)


def kd_pick_dim(var, node_key: int) -> int:
    v = np.ascontiguousarray(var, np.float64)
    return lib().orc_kd_pick_dim(v.ctypes.data_as(_f64p), v.size, node_key)


def partition_injected(X, perm, node_start, node_size, coefs, offsets):
    """Stable side-test partition with given hyperplanes (SolutionKmeans.hpp:189-197)."""
    X, n, d = _points(X)
    perm = np.ascontiguousarray(perm, np.uint32).copy()
    ns = np.ascontiguousarray(node_start, np.uint32)
    nz = np.ascontiguousarray(node_size, np.uint32)
    cf = np.ascontiguousarray(coefs, np.float32)
    of = np.ascontiguousarray(offsets, np.float32)
    lo = np.empty(ns.shape[0], np.uint32)
    lib().orc_partition_injected(_fp(X), n, d, d, _up(perm), ns.shape[0], _up(ns), _up(nz),
                                 _fp(cf), _fp(of), _up(lo))
    return perm, lo


def _state_args(state, n, k):
    if state is None:
        return None, None, None
    ids, dists, sizes = state
    return (np.ascontiguousarray(ids, np.uint32).reshape(n, k),
            np.ascontiguousarray(dists, np.float32).reshape(n, k),
            np.ascontiguousarray(sizes, np.uint32))


def leaf_stage(X, k, perm, leaf_start, leaf_size, state=None, threads: int = 0):
    """addCandidatesLessThan over every leaf (KnnSets.hpp:131-160). State = (ids, dists, sizes)."""
    X, n, d = _points(X)
    ii, di, si = _state_args(state, n, k)
    ids = np.empty((n, k), np.uint32)
    ds = np.empty((n, k), np.float32)
    sz = np.empty(n, np.uint32)
    perm = np.ascontiguousarray(perm, np.uint32)
    ls = np.ascontiguousarray(leaf_start, np.uint32)
    lz = np.ascontiguousarray(leaf_size, np.uint32)
    lib().orc_leaf_stage(_fp(X), n, d, d, k, _up(perm), _up(ls), _up(lz), ls.shape[0], _up(ii),
                         _fp(di), _up(si), _up(ids), _fp(ds), _up(sz), threads)
    return ids, ds, sz


def non(X, k, state, threads: int = 0, rank_lo: int = 0):
    """topUpSingle (SolutionKmeans.hpp:480-555) on a state (ids, dists, sizes); with rank_lo > 0
    topUpSingleOrdered (:393-477), expanding each point's neighbours at ranks >= rank_lo."""
    X, n, d = _points(X)
    ii, di, si = _state_args(state, n, k)
    ids = np.empty((n, k), np.uint32)
    ds = np.empty((n, k), np.float32)
    sz = np.empty(n, np.uint32)
    lib().orc_non(_fp(X), n, d, d, k, _up(ii), _fp(di), _up(si), _up(ids), _fp(ds), _up(sz),
                  threads, rank_lo)
    return ids, ds, sz


def build_graph(X, params: Params, threads: int = 0):
    """constructResult (SolutionKmeans.hpp:558-616). Returns (ids, dists, sizes, stats)."""
    X, n, d = _points(X)
    k = params.k
    ids = np.empty((n, k), np.uint32)
    ds = np.empty((n, k), np.float32)
    sz = np.empty(n, np.uint32)
    st = Stats()
    lib().orc_build(_fp(X), n, d, d, C.byref(params), _up(ids), _fp(ds), _up(sz), threads,
                    C.byref(st))
    return ids, ds, sz, st


def exact_knn(X, queries, k, threads: int = 0):
    """Exact FP64 brute-force k nearest (self excluded), ascending (dist, id)."""
    X, n, d = _points(X)
    q = np.ascontiguousarray(queries, np.uint32)
    ids = np.empty((q.shape[0], k), np.uint32)
    ds = np.empty((q.shape[0], k), np.float64)
    lib().orc_exact_knn(_fp(X), n, d, d, _up(q), q.shape[0], k, _up(ids),
                        ds.ctypes.data_as(_f64p), threads)
    return ids, ds


def recall(result_ids, truth_ids) -> float:
    """Mean |truth ∩ result| / k per row (grade.cpp:61-95 counts; divided by k here)."""
    res = np.asarray(result_ids)
    tru = np.asarray(truth_ids)
    k = tru.shape[1]
    hits = 0
    for r, t in zip(res, tru):
        hits += np.intersect1d(r, t, assume_unique=False).size
    return hits / (k * tru.shape[0])
