"""numpy front end of the CPU oracle (libmvref.so).

TEST INFRASTRUCTURE ONLY: tests, the smoke check and the bench's CPU-baseline
leg use this module as the checker. The product package never imports it.

Each wrapper names the reference routine it restates (paths relative to the
MojoVec v0.8.2 source tree); the C file holds the line-level citations.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmvref.so")
_lib = None

L2 = 0
IP = 1

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int


def build() -> None:
    """Compile libmvref.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "mvref_l2_simd": (ctypes.c_float, [_f32p, _f32p, _i64, _int]),
            "mvref_ip_simd": (ctypes.c_float, [_f32p, _f32p, _i64, _int]),
            "mvref_l2_simd_b4": (ctypes.c_float, [_f32p, _f32p, _i64, _int]),
            "mvref_ip_simd_b4": (ctypes.c_float, [_f32p, _f32p, _i64, _int]),
            "mvref_flat_distances": (None, [_f32p, _int, _int, _f32p, _i64, _i64p, _i64, _int, _f32p]),
            "mvref_flat_pair_distances": (None, [_f32p, _int, _int, _i64p, _i64p, _i64, _f32p]),
            "mvref_sq8_distances": (None, [_vp, _f32p, _i64, _i64p, _i64, _f32p]),
            "mvref_sq8_pair_distances": (None, [_vp, _i64p, _i64p, _i64, _f32p]),
            "mvref_l2_short4": (ctypes.c_float, [_f32p, _f32p, _i64]),
            "mvref_ip_short4": (ctypes.c_float, [_f32p, _f32p, _i64]),
            "mvref_dot_u8": (ctypes.c_uint32, [_u8p, _u8p, _i64]),
            "mvref_dot_i8": (ctypes.c_int32, [_vp, _vp, _i64]),
            "mvref_l2_from_dot": (ctypes.c_uint32, [ctypes.c_uint32] * 3),
            "mvref_l2_simd_batch": (None, [_f32p, _f32p, _i64, _i64, _int, _f32p]),
            "mvref_ip_simd_batch": (None, [_f32p, _f32p, _i64, _i64, _int, _f32p]),
            "mvref_heap_topk_reference": (None, [_f32p, _i64, _int, _f32p, _i64p, ctypes.POINTER(_int)]),
            "mvref_topk_canonical": (None, [_f32p, _vp, _i64, _int, _f32p, _i64p, ctypes.POINTER(_int)]),
            "mvref_philox4x32_10": (None, [ctypes.c_uint64, ctypes.c_uint64, _u32p]),
            "mvref_kmeans_init": (_i64, [_f32p, _i64, _int, _int, ctypes.c_uint64, _f32p]),
            "mvref_kmeans_train": (None, [_f32p, _i64, _int, _int, _int, _f32p, _i64, _int, _f32p, _vp, _int]),
            "mvref_pq_train": (None, [_f32p, _i64, _int, _int, _int, ctypes.c_uint64, _int, _f32p, _int]),
            "mvref_pq_encode": (None, [_f32p, _i64, _int, _int, _f32p, _u8p, _int]),
            "mvref_pq_decode": (None, [_u8p, _i64, _int, _int, _f32p, _f32p]),
            "mvref_pq_lut": (None, [_f32p, _int, _int, _f32p, _int, _f32p]),
            "mvref_flat_search": (None, [_f32p, _i64, _int, _int, _f32p, _i64, _int, _vp, _i64, _f32p, _i64p, _int]),
            "mvref_sq8_encode_vec": (ctypes.c_uint32, [_f32p, _i64, _int, ctypes.c_float, ctypes.c_float, _u8p]),
            "mvref_sq8_create": (_vp, [_int, _int]),
            "mvref_sq8_destroy": (None, [_vp]),
            "mvref_sq8_add": (_int, [_vp, _f32p, _i64]),
            "mvref_sq8_params": (None, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                        ctypes.POINTER(ctypes.c_float), ctypes.POINTER(_i64)]),
            "mvref_sq8_codes": (None, [_vp, _u8p, _u32p]),
            "mvref_sq8_search": (None, [_vp, _f32p, _i64, _int, _int, _vp, _i64, _f32p, _i64p, _int]),
            "mvref_ivfpq_create": (_vp, [_int, _int, _int, _int]),
            "mvref_ivfpq_destroy": (None, [_vp]),
            "mvref_ivfpq_set_codebooks": (None, [_vp, _f32p, _f32p]),
            "mvref_ivfpq_train": (None, [_vp, _f32p, _i64, _int, _int, ctypes.c_uint64, _int]),
            "mvref_ivfpq_get_codebooks": (None, [_vp, _f32p, _f32p]),
            "mvref_ivfpq_add": (None, [_vp, _f32p, _i64, _int, _int]),
            "mvref_ivfpq_add_ids": (None, [_vp, _f32p, _i64, _vp, _int, _int]),
            "mvref_ivfpq_load_lists": (None, [_vp, _i64p, _i32p, _u8p, _vp, _i64]),
            "mvref_ivfpq_list_size": (_i64, [_vp, _int]),
            "mvref_ivfpq_list": (None, [_vp, _int, _i64p, _u8p]),
            "mvref_ivfflat_create": (_vp, [_int, _int, _int, _int]),
            "mvref_ivfflat_destroy": (None, [_vp]),
            "mvref_ivfflat_set_centroids": (None, [_vp, _f32p]),
            "mvref_ivfflat_get_centroids": (None, [_vp, _f32p]),
            "mvref_ivfflat_train": (None, [_vp, _f32p, _i64, _int, ctypes.c_uint64, _int]),
            "mvref_ivfflat_add": (None, [_vp, _f32p, _i64, _vp, _int]),
            "mvref_ivfflat_load_lists": (None, [_vp, _i64p, _i32p, _f32p, _i64]),
            "mvref_ivfflat_list_size": (_i64, [_vp, _int]),
            "mvref_ivfflat_list": (None, [_vp, _int, _i64p, _f32p]),
            "mvref_ivfflat_search": (None, [_vp, _f32p, _i64, _int, _int, _vp, _i64, _f32p, _i64p,
                                            _int]),
            "mvref_ivfpq_search": (None, [_vp, _f32p, _i64, _int, _int, _vp, _int, _vp, _i64,
                                          _f32p, _i64p, _int]),
            "mvref_ivfpq_search_refstruct": (None, [_vp, _f32p, _i64, _int, _int, _int, _f32p, _i64p,
                                                    _int]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


KMEANS_SEED = 0x4D4F4A4F564543  # kmeans.mojo:70


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# ---- distances: mojovec/utils/distances.mojo ----
def l2_simd(x, y, W=8):
    x, y = _f32(x), _f32(y)
    return float(lib().mvref_l2_simd(x, y, x.size, W))


def ip_simd(x, y, W=8):
    x, y = _f32(x), _f32(y)
    return float(lib().mvref_ip_simd(x, y, x.size, W))


def l2_simd_b4(x, y, W=8):
    """One row of l2_distance_simd_batch4[W] (distances.mojo:541-605)."""
    x, y = _f32(x), _f32(y)
    return float(lib().mvref_l2_simd_b4(x, y, x.size, W))


def ip_simd_b4(x, y, W=8):
    """One row of inner_product_simd_batch4[W] (distances.mojo:608-660)."""
    x, y = _f32(x), _f32(y)
    return float(lib().mvref_ip_simd_b4(x, y, x.size, W))


def flat_distances(base, q, ids, metric=L2, order=0):
    """FlatDistanceComputer.distance (order 0) / distance_batch4 (order 1) for
    ids [nq][per_query] (index_flat.mojo:36-115)."""
    base, q = _f32(base), _f32(q)
    ids = np.ascontiguousarray(ids, np.int64)
    nq, per = ids.shape
    out = np.empty((nq, per), np.float32)
    lib().mvref_flat_distances(base, base.shape[1], metric, q, nq, ids, per, order, out)
    return out


def flat_pair_distances(base, a, b, metric=L2):
    """FlatDistanceComputer.symmetric_distance (index_flat.mojo:92-99)."""
    base = _f32(base)
    a, b = np.ascontiguousarray(a, np.int64), np.ascontiguousarray(b, np.int64)
    out = np.empty(a.size, np.float32)
    lib().mvref_flat_pair_distances(base, base.shape[1], metric, a, b, a.size, out)
    return out


def l2_short4(x, y):
    x, y = _f32(x), _f32(y)
    return float(lib().mvref_l2_short4(x, y, x.size))


def ip_short4(x, y):
    x, y = _f32(x), _f32(y)
    return float(lib().mvref_ip_short4(x, y, x.size))


def l2_simd_batch(x, y, W=8):
    x, y = _f32(x), _f32(y)
    out = np.empty(x.shape[0], np.float32)
    lib().mvref_l2_simd_batch(x, y, x.shape[0], x.shape[1], W, out)
    return out


def ip_simd_batch(x, y, W=8):
    x, y = _f32(x), _f32(y)
    out = np.empty(x.shape[0], np.float32)
    lib().mvref_ip_simd_batch(x, y, x.shape[0], x.shape[1], W, out)
    return out


def dot_u8(x, y):
    x = np.ascontiguousarray(x, np.uint8)
    y = np.ascontiguousarray(y, np.uint8)
    return int(lib().mvref_dot_u8(x, y, x.size))


def dot_i8(x, y):
    x = np.ascontiguousarray(x, np.int8)
    y = np.ascontiguousarray(y, np.int8)
    return int(lib().mvref_dot_i8(_ptr(x), _ptr(y), x.size))


def l2_from_dot(na, nb, dot):
    return int(lib().mvref_l2_from_dot(na, nb, dot))


# ---- top-k: mojovec/utils/heap.mojo ----
def heap_topk_reference(dist, k):
    dist = _f32(dist)
    od = np.zeros(k, np.float32)
    oi = np.full(k, -1, np.int64)
    found = ctypes.c_int(0)
    lib().mvref_heap_topk_reference(dist, dist.size, k, od, oi, ctypes.byref(found))
    return od[: found.value], oi[: found.value]


def topk_canonical(dist, k, ids=None):
    dist = _f32(dist)
    ids_a = None if ids is None else np.ascontiguousarray(ids, np.int64)
    od = np.zeros(k, np.float32)
    oi = np.full(k, -1, np.int64)
    found = ctypes.c_int(0)
    lib().mvref_topk_canonical(dist, _ptr(ids_a), dist.size, k, od, oi, ctypes.byref(found))
    return od[: found.value], oi[: found.value]


# ---- k-means: mojovec/clustering/kmeans.mojo ----
def philox(seed, counter):
    out = np.zeros(4, np.uint32)
    lib().mvref_philox4x32_10(seed, counter, out)
    return out


def kmeans_init(x, k, seed=KMEANS_SEED):
    x = _f32(x)
    n, d = x.shape
    c = np.empty((k, d), np.float32)
    off = lib().mvref_kmeans_init(x, n, d, k, seed, c)
    return c, int(off)


def kmeans_train(x, k, niter, init=None, sample_offset=None, fp64_sums=True,
                 seed=KMEANS_SEED, nthreads=0):
    x = _f32(x)
    n, d = x.shape
    if init is None:
        init, sample_offset = kmeans_init(x, k, seed)
    init = _f32(init)
    if sample_offset is None:
        sample_offset = kmeans_init(x, k, seed)[1]
    c = np.empty((k, d), np.float32)
    a = np.empty(n, np.int32)
    lib().mvref_kmeans_train(x, n, d, k, niter, init, sample_offset, int(fp64_sums), c,
                             _ptr(a), nthreads)
    return c, a


# ---- PQ: mojovec/quantization/pq.mojo ----
def pq_train(x, M, niter=15, seed=KMEANS_SEED, fp64_sums=True, nthreads=0):
    x = _f32(x)
    n, d = x.shape
    cb = np.empty((M, 256, d // M), np.float32)
    lib().mvref_pq_train(x, n, d, M, niter, seed, int(fp64_sums), cb, nthreads)
    return cb


def pq_encode(x, codebooks, nthreads=0):
    x = _f32(x)
    n, d = x.shape
    M = codebooks.shape[0]
    codes = np.empty((n, M), np.uint8)
    lib().mvref_pq_encode(x, n, d, M, _f32(codebooks), codes, nthreads)
    return codes


def pq_decode(codes, codebooks):
    codes = np.ascontiguousarray(codes, np.uint8)
    n, M = codes.shape
    d = M * codebooks.shape[2]
    x = np.empty((n, d), np.float32)
    lib().mvref_pq_decode(codes, n, d, M, _f32(codebooks), x)
    return x


def pq_lut(q, codebooks, metric=L2):
    q = _f32(q)
    M = codebooks.shape[0]
    lut = np.empty((M, 256), np.float32)
    lib().mvref_pq_lut(q, q.size, M, _f32(codebooks), metric, lut)
    return lut


# ---- exact Flat: mojovec/index/index_flat.mojo ----
def flat_search(base, q, k, metric=L2, exclude=None, nthreads=0):
    base, q = _f32(base), _f32(q)
    n, d = base.shape
    nq = q.shape[0]
    D = np.empty((nq, k), np.float32)
    I = np.empty((nq, k), np.int64)
    ex = None if exclude is None else np.ascontiguousarray(exclude, np.uint8)
    lib().mvref_flat_search(base, n, d, metric, q, nq, k, _ptr(ex),
                            0 if ex is None else ex.size, D, I, nthreads)
    return D, I


# ---- SQ8: mojovec/quantization/sq8.mojo + index/index_flat_sq8.mojo ----
def sq8_encode(x, symmetric, minimum, inverse_scale):
    x = _f32(x)
    code = np.empty(x.size, np.uint8)
    norm = lib().mvref_sq8_encode_vec(x, x.size, int(symmetric), minimum, inverse_scale, code)
    return (code.view(np.int8) if symmetric else code), int(norm)


def sq8_encode_vec(x, symmetric, minimum, inverse_scale):
    """_encode_sq8_vector (sq8.mojo:13-37): (codes, UInt32 code norm)."""
    x = _f32(x).reshape(-1)
    code = np.empty(x.size, np.uint8)
    n = lib().mvref_sq8_encode_vec(x, x.size, int(bool(symmetric)), float(minimum), float(inverse_scale), code)
    return code, int(n)


class SQ8Index:
    def __init__(self, d, metric=L2):
        self.d = d
        self.metric = metric
        self.h = lib().mvref_sq8_create(d, metric)

    def __del__(self):
        if getattr(self, "h", None):
            lib().mvref_sq8_destroy(self.h)
            self.h = None

    def add(self, x):
        x = _f32(x).reshape(-1, self.d)
        lib().mvref_sq8_add(self.h, x, x.shape[0])

    def params(self):
        a, b, s = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        n = ctypes.c_int64()
        lib().mvref_sq8_params(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(s), ctypes.byref(n))
        return a.value, b.value, s.value, n.value

    def codes(self):
        n = self.params()[3]
        c = np.empty((n, self.d), np.uint8)
        nr = np.empty(n, np.uint32)
        lib().mvref_sq8_codes(self.h, c, nr)
        return c, nr

    def distances(self, q, ids):
        """SQ8DistanceComputer.distance for ids [nq][per_query]."""
        q = _f32(q).reshape(-1, self.d)
        ids = np.ascontiguousarray(ids, np.int64)
        out = np.empty(ids.shape, np.float32)
        lib().mvref_sq8_distances(self.h, q, ids.shape[0], ids, ids.shape[1], out)
        return out

    def pair_distances(self, a, b):
        """SQ8DistanceComputer.symmetric_distance over row pairs."""
        a, b = np.ascontiguousarray(a, np.int64), np.ascontiguousarray(b, np.int64)
        out = np.empty(a.size, np.float32)
        lib().mvref_sq8_pair_distances(self.h, a, b, a.size, out)
        return out

    def search(self, q, k, k_factor=1, exclude=None, nthreads=0):
        q = _f32(q).reshape(-1, self.d)
        nq = q.shape[0]
        D = np.empty((nq, k), np.float32)
        I = np.empty((nq, k), np.int64)
        ex = None if exclude is None else np.ascontiguousarray(exclude, np.uint8)
        lib().mvref_sq8_search(self.h, q, nq, k, k_factor, _ptr(ex), 0 if ex is None else ex.size,
                               D, I, nthreads)
        return D, I


# ---- IVF-PQ: mojovec/index/index_ivf_pq.mojo ----
class IVFPQIndex:
    def __init__(self, d, nlist, M, metric=L2):
        self.d, self.nlist, self.M, self.metric = d, nlist, M, metric
        self.h = lib().mvref_ivfpq_create(d, nlist, M, metric)

    def __del__(self):
        if getattr(self, "h", None):
            lib().mvref_ivfpq_destroy(self.h)
            self.h = None

    def set_codebooks(self, coarse, pq):
        lib().mvref_ivfpq_set_codebooks(self.h, _f32(coarse), _f32(pq))

    def train(self, x, niter_coarse=15, niter_pq=15, seed=KMEANS_SEED, nthreads=0):
        x = _f32(x).reshape(-1, self.d)
        lib().mvref_ivfpq_train(self.h, x, x.shape[0], niter_coarse, niter_pq, seed, nthreads)

    def codebooks(self):
        c = np.empty((self.nlist, self.d), np.float32)
        p = np.empty((self.M, 256, self.d // self.M), np.float32)
        lib().mvref_ivfpq_get_codebooks(self.h, c, p)
        return c, p

    def add(self, x, keep_vectors=True, nthreads=0):
        x = _f32(x).reshape(-1, self.d)
        lib().mvref_ivfpq_add(self.h, x, x.shape[0], int(keep_vectors), nthreads)

    def add_with_ids(self, x, ids, keep_vectors=True, nthreads=0):
        """IndexIVFPQ.add_with_ids (index_ivf_pq.mojo:150-213)."""
        x = _f32(x).reshape(-1, self.d)
        ids = np.ascontiguousarray(ids, np.int64).reshape(-1)
        lib().mvref_ivfpq_add_ids(self.h, x, x.shape[0], _ptr(ids), int(keep_vectors), nthreads)

    def load_lists(self, list_off, ids, codes, vectors=None):
        list_off = np.ascontiguousarray(list_off, np.int64)
        vec = None if vectors is None else _f32(vectors)
        lib().mvref_ivfpq_load_lists(self.h, list_off, np.ascontiguousarray(ids, np.int32),
                                     np.ascontiguousarray(codes, np.uint8).reshape(-1), _ptr(vec),
                                     int(list_off[-1]))

    def list(self, l):
        n = lib().mvref_ivfpq_list_size(self.h, l)
        ids = np.empty(n, np.int64)
        codes = np.empty((n, self.M), np.uint8)
        lib().mvref_ivfpq_list(self.h, l, ids, codes)
        return ids, codes

    def search(self, q, k, nprobe, k_factor=1, nprobe_per_query=None, exclude=None, nthreads=0):
        q = _f32(q).reshape(-1, self.d)
        nq = q.shape[0]
        D = np.empty((nq, k), np.float32)
        I = np.empty((nq, k), np.int64)
        npq = None if nprobe_per_query is None else np.ascontiguousarray(nprobe_per_query, np.int32)
        ex = None if exclude is None else np.ascontiguousarray(exclude, np.uint8)
        lib().mvref_ivfpq_search(self.h, q, nq, k, nprobe, _ptr(npq), k_factor, _ptr(ex),
                                 0 if ex is None else ex.size, D, I, nthreads)
        return D, I

    def search_refstruct(self, q, k, nprobe, k_factor=1, nthreads=0):
        """The search with the reference's threading (index_ivf_pq.mojo:
        256-276): coarse step parallel, LUT + ADC loop serial."""
        q = _f32(q).reshape(-1, self.d)
        nq = q.shape[0]
        D = np.empty((nq, k), np.float32)
        I = np.empty((nq, k), np.int64)
        lib().mvref_ivfpq_search_refstruct(self.h, q, nq, k, nprobe, k_factor, D, I, nthreads)
        return D, I


# ---- IVF-Flat: mojovec/index/index_ivf_flat.mojo ----
class IVFFlatIndex:
    def __init__(self, d, nlist, metric=L2, quantizer_metric=None):
        self.d, self.nlist, self.metric = d, nlist, metric
        qm = metric if quantizer_metric is None else quantizer_metric
        self.h = lib().mvref_ivfflat_create(d, nlist, metric, qm)

    def __del__(self):
        if getattr(self, "h", None):
            lib().mvref_ivfflat_destroy(self.h)
            self.h = None

    def set_centroids(self, coarse):
        lib().mvref_ivfflat_set_centroids(self.h, _f32(coarse))

    def centroids(self):
        c = np.empty((self.nlist, self.d), np.float32)
        lib().mvref_ivfflat_get_centroids(self.h, c)
        return c

    def train(self, x, niter=15, seed=KMEANS_SEED, nthreads=0):
        x = _f32(x).reshape(-1, self.d)
        lib().mvref_ivfflat_train(self.h, x, x.shape[0], niter, seed, nthreads)

    def add(self, x, ids=None, nthreads=0):
        x = _f32(x).reshape(-1, self.d)
        idv = None if ids is None else np.ascontiguousarray(ids, np.int64)
        lib().mvref_ivfflat_add(self.h, x, x.shape[0], _ptr(idv), nthreads)

    def load_lists(self, list_off, ids, rows):
        list_off = np.ascontiguousarray(list_off, np.int64)
        lib().mvref_ivfflat_load_lists(self.h, list_off, np.ascontiguousarray(ids, np.int32),
                                       _f32(rows).reshape(-1), int(list_off[-1]))

    def list(self, l):
        n = lib().mvref_ivfflat_list_size(self.h, l)
        ids = np.empty(n, np.int64)
        rows = np.empty((n, self.d), np.float32)
        lib().mvref_ivfflat_list(self.h, l, ids, rows)
        return ids, rows

    def search(self, q, k, nprobe, exclude=None, nthreads=0):
        q = _f32(q).reshape(-1, self.d)
        nq = q.shape[0]
        D = np.empty((nq, k), np.float32)
        I = np.empty((nq, k), np.int64)
        ex = None if exclude is None else np.ascontiguousarray(exclude, np.uint8)
        lib().mvref_ivfflat_search(self.h, q, nq, k, nprobe, _ptr(ex), 0 if ex is None else ex.size,
                                   D, I, nthreads)
        return D, I
