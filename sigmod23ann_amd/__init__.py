"""MI355X (gfx950) approximate KNN-graph builder: Python view of ``libknng.so``.

The product is the C ABI in ``include/knng.h`` (HIP kernels + C++ host driver in ``csrc/``).
This module loads it with ctypes and mirrors the reference's entry point::

    construct_result(points, k=100, ...)  ->  (n, k) uint32 ids

which replaces ``SolutionKmeans::constructResult`` (src/SolutionKmeans.hpp:558-616). There is no
CPU fallback: every call goes to the GPU library, and importing fails loudly when the library has
not been built.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libknng.so")
CLI_PATH = os.path.join(_HERE, "knng_cli")      # src/main.cpp's flow over the C ABI
GRADE_PATH = os.path.join(_HERE, "knng_grade")  # src/grade.cpp's semantics
CSRC = os.path.join(_HERE, "csrc")

OK = 0
ERR_INVALID = -1
ERR_HIP = -2
ERR_UNSUPPORTED = -3
ERR_NOMEM = -4
ERR_OVERFLOW = -5

PREC_FP32 = 0
PREC_FP16 = 1
PREC_BF16 = 2

FLAG_DIRECT_NON = 1
FLAG_BIG_LEAF = 2
FLAG_INCREMENTAL_NON = 4
FLAG_SMALL_NON_CAP = 8  # tests: tiny candidate buffers, so the overflow path runs
FLAG_NON_SCREEN16 = 64  # neighbours-of-neighbours screen on a 16-bit Gram of rows centred on q
FLAG_NON_FP32_GRAM = 512  # neighbours-of-neighbours screen on the FP32 Gram (not the fp16 one)
FLAG_LEAF_BATCHED = 128  # leaf stage: one persistent kernel over the round's work list
FLAG_LEAF_GRAPH = 256    # ... replayed from a captured hipGraph
FLAG_KD_SPLIT = 16  # random-rotation KD splitter (SolutionRandomKD) instead of 2-means
FLAG_AUTOTUNE = 32  # leaf workgroup shape picked by timing both on the first launch of a size bin
FLAG_NO_BUCKET_LIMIT = 1024  # exact form: leaf screen without the in-bucket k-th upper-bound limit
FLAG_LEAF_FP32_GRAM = 2048  # exact form: leaf screen on the FP32 Gram, every bucket on the big kernel
FLAG_NON_ONE_SET = 4096  # tests: one set of NoN candidate buffers (as a build near device capacity)
FLAG_SIDE_FP32 = 8192  # split side test on the FP32 rows only (no fp16 copy with its exact bound)

GEN_LOWRANK = 0
GEN_MIXTURE = 1
GEN_ZIPF = 2

# Every symbol include/knng.h declares.
EXPORTS = [
    "knng_last_error", "knng_default_params", "knng_create", "knng_destroy", "knng_set_params",
    "knng_get_stats", "knng_build", "knng_build_device", "knng_split_round",
    "knng_partition_injected", "knng_leaf_stage", "knng_leaf_stage_device_noalloc", "knng_non", "knng_exact_knn_device",
    "knng_exact_upper_from_lists", "knng_recall", "knng_generate_device", "knng_device_alloc",
    "knng_device_free", "knng_memcpy_h2d", "knng_memcpy_d2h", "knng_memcpy_d2d",
    "knng_synchronize", "knng_mem_get_info", "knng_job_footprint", "knng_server_create", "knng_server_submit",
    "knng_server_poll", "knng_server_wait", "knng_server_get_stats", "knng_server_destroy",
]


def build(force: bool = True, jobs: int = 8) -> str:
    """Compile libknng.so for gfx950 with hipcc (csrc/Makefile)."""
    args = ["make", "-C", CSRC, f"-j{jobs}"]
    if force:
        args.append("-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)
    return LIB_PATH


class Params(C.Structure):
    _fields_ = [
        ("k", C.c_uint32),
        ("leaf_max", C.c_uint32),
        ("leaf_max_odd", C.c_uint32),
        ("small_max", C.c_uint32),
        ("iterations", C.c_uint32),
        ("rounds", C.c_uint32),
        ("non_passes", C.c_uint32),
        ("precision", C.c_uint32),
        ("rerank", C.c_uint32),
        ("flags", C.c_uint32),
        ("search_k", C.c_uint32),
        ("non_rank_lo", C.c_uint32),
        ("seed", C.c_uint64),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("split_ms", C.c_double),
        ("side_ms", C.c_double),
        ("leaf_ms", C.c_double),
        ("non_ms", C.c_double),
        ("total_ms", C.c_double),
        ("side_launches", C.c_uint64),
        ("side_rows", C.c_uint64),
        ("leaf_launches", C.c_uint64),
        ("leaves", C.c_uint64),
        ("leaf_pairs", C.c_uint64),
        ("survivors", C.c_uint64),
        ("inserts", C.c_uint64),
        ("split_failures", C.c_uint64),
        ("split_fallbacks", C.c_uint64),
        ("non_overflow", C.c_uint64),
        ("non_screen_ms", C.c_double),
        ("non_screen_launches", C.c_uint64),
        ("non_screen_bytes", C.c_uint64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Job(C.Structure):
    _fields_ = [
        ("points", C.c_void_p),
        ("n", C.c_uint32),
        ("d", C.c_uint32),
        ("row_stride", C.c_size_t),
        ("points_on_device", C.c_uint32),
        ("params", Params),
        ("out_ids", C.c_void_p),
        ("out_dists", C.c_void_p),
    ]


class JobResult(C.Structure):
    _fields_ = [
        ("status", C.c_int),
        ("queue_ms", C.c_double),
        ("run_ms", C.c_double),
        ("latency_ms", C.c_double),
        ("footprint", C.c_uint64),
        ("worker", C.c_uint32),
        ("start_ms", C.c_double),
        ("done_ms", C.c_double),
        ("worker_device_bytes", C.c_uint64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class ServerStats(C.Structure):
    _fields_ = [
        ("budget", C.c_uint64),
        ("in_use", C.c_uint64),
        ("peak_in_use", C.c_uint64),
        ("trims", C.c_uint64),
        ("queued", C.c_uint64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class KnngError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"knng error {code}: {msg}")
        self.code = code


_lib = None
_P = C.POINTER
_u32p = _P(C.c_uint32)
_f32p = _P(C.c_float)
_f64p = _P(C.c_double)
_vp = C.c_void_p
_sz = C.c_size_t
_u32 = C.c_uint32


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with sigmod23ann_amd.build()")
        L = C.CDLL(LIB_PATH)
        L.knng_last_error.restype = C.c_char_p
        L.knng_default_params.argtypes = [_P(Params)]
        L.knng_create.argtypes = [_P(Params), _P(_vp)]
        L.knng_destroy.argtypes = [_vp]
        L.knng_set_params.argtypes = [_vp, _P(Params)]
        L.knng_get_stats.argtypes = [_vp, _P(Stats)]
        L.knng_build.argtypes = [_vp, _f32p, _u32, _u32, _sz, _u32p, _f32p]
        L.knng_build_device.argtypes = [_vp, _vp, _u32, _u32, _sz, _vp, _vp]
        L.knng_split_round.argtypes = [_vp, _f32p, _u32, _u32, _sz, _u32, _u32p, _u32p, _u32p,
                                       _u32p]
        L.knng_partition_injected.argtypes = [_vp, _f32p, _u32, _u32, _sz, _u32p, _u32, _u32p,
                                              _u32p, _f32p, _f32p, _u32p]
        L.knng_leaf_stage.argtypes = [_vp, _f32p, _u32, _u32, _sz, _u32p, _u32p, _u32p, _u32,
                                      _u32p, _f32p, C.c_int]
        L.knng_leaf_stage_device_noalloc.argtypes = [_vp, _vp, _u32, _u32, _sz, _vp, _u32p, _u32p,
                                                     _u32, _vp, _vp, C.c_int]
        L.knng_non.argtypes = [_vp, _f32p, _u32, _u32, _sz, _u32p, _f32p]
        L.knng_exact_knn_device.argtypes = [_vp, _vp, _u32, _u32, _sz, _vp, _u32, _u32, _vp,
                                            _u32, _vp, _vp]
        L.knng_exact_upper_from_lists.argtypes = [_vp, _vp, _u32, _u32, _sz, _vp, _u32, _vp,
                                                  _u32, _vp]
        L.knng_recall.argtypes = [_u32p, _u32p, _u32, _u32]
        L.knng_recall.restype = C.c_double
        L.knng_generate_device.argtypes = [_vp, C.c_int, _u32, _u32, _sz, C.c_uint64, _vp]
        L.knng_device_alloc.argtypes = [_sz, _P(_vp)]
        L.knng_device_free.argtypes = [_vp]
        L.knng_memcpy_h2d.argtypes = [_vp, _vp, _sz]
        L.knng_memcpy_d2h.argtypes = [_vp, _vp, _sz]
        L.knng_memcpy_d2d.argtypes = [_vp, _vp, _sz]
        L.knng_job_footprint.argtypes = [_u32, _u32, _P(Params)]
        L.knng_job_footprint.restype = C.c_uint64
        L.knng_server_create.argtypes = [_u32, C.c_uint64, _P(_vp)]
        L.knng_server_submit.argtypes = [_vp, _P(Job), _P(C.c_uint64)]
        L.knng_server_poll.argtypes = [_vp, C.c_uint64, _P(C.c_int)]
        L.knng_server_wait.argtypes = [_vp, C.c_uint64, _P(JobResult)]
        L.knng_server_get_stats.argtypes = [_vp, _P(ServerStats)]
        L.knng_mem_get_info.argtypes = [_P(C.c_size_t), _P(C.c_size_t)]
        L.knng_server_destroy.argtypes = [_vp]
        L.knng_server_destroy.restype = None
        _lib = L
    return _lib


def _check(code):
    if code != OK:
        raise KnngError(code, lib().knng_last_error().decode(errors="replace"))


def params(k=100, leaf_max=400, leaf_max_odd=0, small_max=3000, iterations=1, rounds=6,
           non_passes=2, precision=PREC_FP32, rerank=160, seed=12345, flags=0, search_k=0,
           non_rank_lo=0) -> Params:
    return Params(k, leaf_max, leaf_max_odd, small_max, iterations, rounds, non_passes, precision,
                  rerank, flags, search_k, non_rank_lo, seed)


def _f(a):
    return a.ctypes.data_as(_f32p)


def _u(a):
    return a.ctypes.data_as(_u32p)


def _points(X):
    X = np.ascontiguousarray(X, dtype=np.float32)
    if X.ndim != 2:
        raise ValueError("points must be 2-D (n, d)")
    return X, X.shape[0], X.shape[1]


class DeviceArray:
    """Device allocation owned by Python (plain hipMalloc through the library)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = _vp()
        _check(lib().knng_device_alloc(self.nbytes, C.byref(p)))
        self.ptr = p.value

    def free(self):
        if getattr(self, "ptr", None):
            lib().knng_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def upload(self, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _check(lib().knng_memcpy_h2d(self.ptr, arr.ctypes.data, arr.nbytes))

    def download(self, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        _check(lib().knng_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out


class Builder:
    """One handle: device buffers and a HIP stream reused across calls (not reentrant)."""

    def __init__(self, prm: Params | None = None, **kw):
        self.prm = prm if prm is not None else params(**kw)
        h = _vp()
        _check(lib().knng_create(C.byref(self.prm), C.byref(h)))
        self.h = h.value

    def close(self):
        if getattr(self, "h", None):
            lib().knng_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, prm: Params):
        _check(lib().knng_set_params(self.h, C.byref(prm)))
        self.prm = prm

    def stats(self) -> dict:
        s = Stats()
        _check(lib().knng_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    # -- whole build (constructResult) --------------------------------------------------------
    def build(self, X, return_dists=False):
        X, n, d = _points(X)
        k = self.prm.k
        ids = np.empty((n, k), np.uint32)
        ds = np.empty((n, k), np.float32) if return_dists else None
        _check(lib().knng_build(self.h, _f(X), n, d, d, _u(ids), _f(ds) if ds is not None else None))
        return (ids, ds) if return_dists else ids

    def build_device(self, d_points: int, n: int, d: int, row_stride: int, d_out_ids: int,
                     d_out_dists: int | None = None):
        _check(lib().knng_build_device(self.h, d_points, n, d, row_stride, d_out_ids,
                                       d_out_dists))

    # -- stage entry points -------------------------------------------------------------------
    def split_round(self, X, round_idx: int):
        X, n, d = _points(X)
        perm = np.empty(n, np.uint32)
        ls = np.empty(max(n, 1), np.uint32)
        lz = np.empty(max(n, 1), np.uint32)
        nl = C.c_uint32(0)
        _check(lib().knng_split_round(self.h, _f(X), n, d, d, round_idx, _u(perm), _u(ls), _u(lz),
                                      C.byref(nl)))
        return perm, ls[: nl.value].copy(), lz[: nl.value].copy()

    def partition_injected(self, X, perm, node_start, node_size, coefs, offsets):
        X, n, d = _points(X)
        perm = np.ascontiguousarray(perm, np.uint32).copy()
        ns = np.ascontiguousarray(node_start, np.uint32)
        nz = np.ascontiguousarray(node_size, np.uint32)
        cf = np.ascontiguousarray(coefs, np.float32)
        of = np.ascontiguousarray(offsets, np.float32)
        lo = np.empty(ns.shape[0], np.uint32)
        _check(lib().knng_partition_injected(self.h, _f(X), n, d, d, _u(perm), ns.shape[0], _u(ns),
                                             _u(nz), _f(cf), _f(of), _u(lo)))
        return perm, lo

    def leaf_stage(self, X, perm, leaf_start, leaf_size, state=None):
        """state = (ids, dists) each (n, k), rows sorted by (dist, id); None = empty sets."""
        X, n, d = _points(X)
        k = self.prm.k
        if state is None:
            ids = np.empty((n, k), np.uint32)
            ds = np.empty((n, k), np.float32)
            empty = 1
        else:
            ids = np.ascontiguousarray(state[0], np.uint32).copy()
            ds = np.ascontiguousarray(state[1], np.float32).copy()
            empty = 0
        perm = np.ascontiguousarray(perm, np.uint32)
        ls = np.ascontiguousarray(leaf_start, np.uint32)
        lz = np.ascontiguousarray(leaf_size, np.uint32)
        _check(lib().knng_leaf_stage(self.h, _f(X), n, d, d, _u(perm), _u(ls), _u(lz), ls.shape[0],
                                     _u(ids), _f(ds), empty))
        return ids, ds

    def leaf_stage_device_noalloc(self, d_points, n, d, row_stride, d_perm, leaf_start,
                                  leaf_size, d_ids, d_dists, start_empty):
        """knng_leaf_stage on device points, permutation and lists, with the library's fixed
        scratch instead of allocated buffers (leaf tables on the host)."""
        ls = np.ascontiguousarray(leaf_start, np.uint32)
        lz = np.ascontiguousarray(leaf_size, np.uint32)
        _check(lib().knng_leaf_stage_device_noalloc(self.h, d_points, n, d, row_stride, d_perm,
                                                    _u(ls), _u(lz), ls.shape[0], d_ids, d_dists,
                                                    1 if start_empty else 0))

    def non(self, X, state):
        X, n, d = _points(X)
        ids = np.ascontiguousarray(state[0], np.uint32).copy()
        ds = np.ascontiguousarray(state[1], np.float32).copy()
        _check(lib().knng_non(self.h, _f(X), n, d, d, _u(ids), _f(ds)))
        return ids, ds

    # -- recall tooling -------------------------------------------------------------------------
    def exact_knn_device(self, d_points, n, d, row_stride, d_queries, nq, k, d_upper, cap,
                         d_out_ids, d_out_dists):
        _check(lib().knng_exact_knn_device(self.h, d_points, n, d, row_stride, d_queries, nq, k,
                                           d_upper, cap, d_out_ids, d_out_dists))

    def exact_upper_from_lists(self, d_points, n, d, row_stride, d_queries, nq, d_lists, k,
                               d_upper):
        _check(lib().knng_exact_upper_from_lists(self.h, d_points, n, d, row_stride, d_queries,
                                                 nq, d_lists, k, d_upper))

    def exact_knn(self, X, queries, k, cap=None):
        """Host convenience: exact FP64 k-NN of query ids (uploads X)."""
        X, n, d = _points(X)
        q = np.ascontiguousarray(queries, np.uint32)
        cap = cap or max(n, k)
        dX = DeviceArray(X.nbytes)
        dX.upload(X)
        dq = DeviceArray(q.nbytes)
        dq.upload(q)
        di = DeviceArray(q.shape[0] * k * 4)
        dd = DeviceArray(q.shape[0] * k * 8)
        self.exact_knn_device(dX.ptr, n, d, d, dq.ptr, q.shape[0], k, None, cap, di.ptr, dd.ptr)
        return di.download((q.shape[0], k), np.uint32), dd.download((q.shape[0], k), np.float64)

    def generate_device(self, kind: int, n: int, d: int, row_stride: int, seed: int,
                        d_points: int):
        _check(lib().knng_generate_device(self.h, kind, n, d, row_stride, seed, d_points))


def mem_get_info():
    """(free, total) device bytes of the current device."""
    f, t = C.c_size_t(), C.c_size_t()
    _check(lib().knng_mem_get_info(C.byref(f), C.byref(t)))
    return f.value, t.value


def recall(result_ids, truth_ids) -> float:
    """grade.cpp:61-95: mean |truth ∩ result| per row, divided by k."""
    r = np.ascontiguousarray(result_ids, np.uint32)
    t = np.ascontiguousarray(truth_ids, np.uint32)
    return lib().knng_recall(_u(r), _u(t), t.shape[0], t.shape[1])


def construct_result(points, k=100, **kw):
    """constructResult (src/SolutionKmeans.hpp:558): (n, d) float32 -> (n, k) uint32 ids, each row
    ascending by distance, short rows padded with id 1."""
    b = Builder(params(k=k, **kw))
    try:
        return b.build(points)
    finally:
        b.close()


def job_footprint(n, d, prm: Params) -> int:
    """Device bytes the server's admission accounts for a job."""
    return int(lib().knng_job_footprint(n, d, C.byref(prm)))


class Ticket:
    """A submitted job's id. The server, not the ticket, holds the job's buffers (the host points
    and the output arrays the worker writes), from submit until wait() returns or close() has
    drained the job, so dropping a ticket early frees nothing the worker still uses."""

    def __init__(self, jid):
        self.id = jid


class Server:
    """Continuous batching of build jobs (knng_server_*): `workers` builders on their own HIP
    streams, jobs admitted in arrival order while the device memory accounted to the workers (what
    each keeps between jobs plus the running jobs' footprints) fits `hbm_budget`."""

    def __init__(self, workers=4, hbm_budget=250 << 30):
        h = _vp()
        _check(lib().knng_server_create(workers, hbm_budget, C.byref(h)))
        self.h = h.value
        self._jobs = {}   # job id -> (ids, dists, kept input); released by wait() or close()

    def submit(self, prm: Params, X=None, device_ptr=None, n=None, d=None, row_stride=None,
               return_dists=False) -> Ticket:
        """Host points X (n, d), or device rows at device_ptr (n rows of row_stride floats; the
        caller keeps that memory valid until the job is waited for or the server closed)."""
        if X is not None:
            X, n, d = _points(X)
            ptr, on_dev, stride, keep = X.ctypes.data, 0, d, X
        else:
            ptr, on_dev, stride, keep = device_ptr, 1, row_stride or d, None
        ids = np.empty((n, prm.k), np.uint32)
        ds = np.empty((n, prm.k), np.float32) if return_dists else None
        job = Job(ptr, n, d, stride, on_dev, prm, ids.ctypes.data,
                  ds.ctypes.data if ds is not None else None)
        jid = C.c_uint64()
        _check(lib().knng_server_submit(self.h, C.byref(job), C.byref(jid)))
        self._jobs[jid.value] = (ids, ds, keep)
        return Ticket(jid.value)

    def done(self, t: Ticket) -> bool:
        flag = C.c_int()
        _check(lib().knng_server_poll(self.h, t.id, C.byref(flag)))
        return bool(flag.value)

    def wait(self, t: Ticket):
        """(ids, dists or None, result dict); raises if the job failed."""
        r = JobResult()
        code = lib().knng_server_wait(self.h, t.id, C.byref(r))
        ids, ds, _ = self._jobs.pop(t.id, (None, None, None))
        _check(code)
        return ids, ds, r.as_dict()

    def stats(self) -> dict:
        st = ServerStats()
        _check(lib().knng_server_get_stats(self.h, C.byref(st)))
        return st.as_dict()

    def close(self):
        """Runs every submitted job to completion, stops the workers, then releases the buffers of
        jobs nobody waited for."""
        if getattr(self, "h", None):
            lib().knng_server_destroy(self.h)
            self.h = None
        self._jobs = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- contest file formats (src/io.h) ---------------------------------------------------------
def read_points(path, dims=100) -> np.ndarray:
    """uint32 N, then N rows of `dims` float32 (ReadBinArray, src/io.h:66-94)."""
    with open(path, "rb") as f:
        n = int(np.frombuffer(f.read(4), np.uint32)[0])
        X = np.frombuffer(f.read(n * dims * 4), np.float32)
    if X.size != n * dims:
        raise ValueError(f"{path}: header says {n} points, file holds {X.size // dims}")
    return X.reshape(n, dims)


def write_points(path, X):
    X = np.ascontiguousarray(X, np.float32)
    with open(path, "wb") as f:
        f.write(np.uint32(X.shape[0]).tobytes())
        f.write(X.tobytes())


def save_graph(path, ids):
    """n rows of k uint32 ids, no header (SaveKNNG, src/io.h:21-33)."""
    np.ascontiguousarray(ids, np.uint32).tofile(path)


def read_graph(path, n, k=100) -> np.ndarray:
    return np.fromfile(path, np.uint32, count=n * k).reshape(n, k)
