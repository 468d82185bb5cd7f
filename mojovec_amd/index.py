"""Index-level wrappers over the C ABI (the reference's Mojo index layer).

Mirrors of mojovec/index/index_flat.mojo (IndexFlat), index_flat_sq8.mojo
(IndexFlatSQ8), index_ivf_pq.mojo (IndexIVFPQ), clustering/kmeans.mojo (KMeans)
and quantization/pq.mojo (ProductQuantizer). Arrays are numpy; results come
back as (distances float32 [nq, k], labels int64 [nq, k]).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib

METRIC_L2 = 0
METRIC_INNER_PRODUCT = 1
KMEANS_SEED = 0x4D4F4A4F564543  # kmeans.mojo:70


def _f32(x, d=None):
    a = np.ascontiguousarray(x, dtype=np.float32)
    if d is not None:
        a = a.reshape(-1, d)
    return a


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _mask(exclude):
    if exclude is None:
        return None, 0
    m = np.ascontiguousarray(exclude, dtype=np.uint8)
    return m, m.size


def _ids2(ids, nq):
    ids = np.ascontiguousarray(ids, np.int64)
    if ids.ndim == 1 and nq == 1:
        ids = ids.reshape(1, -1)
    if ids.ndim != 2 or ids.shape[0] != nq:
        raise ValueError("ids must have shape [nq, per_query]")
    return ids


def _pairs(a, b):
    a = np.ascontiguousarray(a, np.int64).reshape(-1)
    b = np.ascontiguousarray(b, np.int64).reshape(-1)
    if a.size != b.size:
        raise ValueError("a and b must have the same length")
    return a, b


class IndexFlat:
    """Exact brute-force index (IndexFlat, index_flat.mojo:117-330)."""

    def __init__(self, d: int, metric: int = METRIC_L2):
        L = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(L.mvg_flat_create(d, metric, ctypes.byref(h)))
        self._h, self.d, self.metric_type = h, d, metric

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.load().mvg_flat_destroy(self._h)
            self._h = None

    @property
    def ntotal(self) -> int:
        return int(_lib.load().mvg_flat_ntotal(self._h))

    def set_search_mode(self, precision: str = "plain", k_factor: int = 1) -> None:
        """"plain" FP32 (reference order), "mixed" / "mixed-bf16" (BF16 MFMA
        candidates) or "mixed-fp8" (FP8 E4M3 MFMA candidates); mixed forms
        rerank k * k_factor candidates exactly in FP32 when k_factor > 1."""
        p = {"plain": 0, "mixed": 1, "mixed-bf16": 1, "mixed-fp8": 2}[precision]
        _lib.check(_lib.load().mvg_flat_set_search_mode(self._h, p, int(k_factor)))

    def add(self, x) -> None:
        x = _f32(x, self.d)
        _lib.check(_lib.load().mvg_flat_add(self._h, _ptr(x), x.shape[0]))

    def search(self, x, k: int, exclude=None):
        x = _f32(x, self.d)
        nq = x.shape[0]
        D = np.empty((nq, k), np.float32)
        I = np.empty((nq, k), np.int64)
        m, ml = _mask(exclude)
        _lib.check(_lib.load().mvg_flat_search(self._h, _ptr(x), nq, k, _ptr(m), ml, _ptr(D), _ptr(I)))
        return D, I

    def distances(self, x, ids, order: str = "distance"):
        """FlatDistanceComputer distances (index_flat.mojo:36-115) of each
        query row to the rows ids[i] ([nq][per_query]): "distance" or
        "batch4" (distance_batch4) summation order."""
        x = _f32(x, self.d)
        ids = _ids2(ids, x.shape[0])
        o = {"distance": 0, "batch4": 1}[order]
        out = np.empty(ids.shape, np.float32)
        _lib.check(_lib.load().mvg_flat_distances(self._h, _ptr(x), ids.shape[0], _ptr(ids), ids.shape[1], o,
                                                  _ptr(out)))
        return out

    def pair_distances(self, a, b):
        """FlatDistanceComputer.symmetric_distance for row pairs (a[p], b[p])."""
        a, b = _pairs(a, b)
        out = np.empty(a.size, np.float32)
        _lib.check(_lib.load().mvg_flat_pair_distances(self._h, _ptr(a), _ptr(b), a.size, _ptr(out)))
        return out

    def get_vector(self, i: int):
        """IndexFlat.get_vector_span (index_flat.mojo:212-227): row i."""
        out = np.empty(self.d, np.float32)
        _lib.check(_lib.load().mvg_flat_get_vector(self._h, int(i), _ptr(out)))
        return out

    def get_vectors(self, ids):
        """Stored rows ids[0..n) as [n, d]."""
        ids = np.ascontiguousarray(ids, np.int64).reshape(-1)
        out = np.empty((ids.size, self.d), np.float32)
        _lib.check(_lib.load().mvg_flat_get_vectors(self._h, _ptr(ids), ids.size, _ptr(out)))
        return out


class IndexFlatSQ8:
    """Global-calibration SQ8 index (IndexFlatSQ8, index_flat_sq8.mojo:192-582)."""

    def __init__(self, d: int, metric: int = METRIC_L2):
        L = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(L.mvg_sq8_create(d, metric, ctypes.byref(h)))
        self._h, self.d, self.metric_type = h, d, metric

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.load().mvg_sq8_destroy(self._h)
            self._h = None

    def add(self, x) -> None:
        x = _f32(x, self.d)
        _lib.check(_lib.load().mvg_sq8_add(self._h, _ptr(x), x.shape[0]))

    def params(self):
        a, b, s = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        n = ctypes.c_int64()
        _lib.check(_lib.load().mvg_sq8_params(self._h, ctypes.byref(a), ctypes.byref(b),
                                               ctypes.byref(s), ctypes.byref(n)))
        return a.value, b.value, s.value, n.value

    @property
    def ntotal(self) -> int:
        return self.params()[3]

    def set_search_mode(self, use_mfma: bool = True) -> None:
        """int8 MFMA code contraction (default) or dot4 VALU; same results."""
        _lib.check(_lib.load().mvg_sq8_set_search_mode(self._h, int(use_mfma)))

    @property
    def scale(self) -> float:
        return self.params()[2]

    def codes(self):
        n = self.ntotal
        c = np.empty((n, self.d), np.uint8)
        nr = np.empty(n, np.uint32)
        _lib.check(_lib.load().mvg_sq8_get_codes(self._h, _ptr(c), _ptr(nr)))
        return c, nr

    def distances(self, x, ids):
        """SQ8DistanceComputer distances (index_flat_sq8.mojo:21-190) of each
        query row to the rows ids[i] ([nq][per_query])."""
        x = _f32(x, self.d)
        ids = _ids2(ids, x.shape[0])
        out = np.empty(ids.shape, np.float32)
        _lib.check(_lib.load().mvg_sq8_distances(self._h, _ptr(x), ids.shape[0], _ptr(ids), ids.shape[1],
                                                 _ptr(out)))
        return out

    def pair_distances(self, a, b):
        """SQ8DistanceComputer.symmetric_distance for row pairs (a[p], b[p])."""
        a, b = _pairs(a, b)
        out = np.empty(a.size, np.float32)
        _lib.check(_lib.load().mvg_sq8_pair_distances(self._h, _ptr(a), _ptr(b), a.size, _ptr(out)))
        return out

    def get_vector(self, i: int):
        """IndexFlatSQ8.get_vector_span (index_flat_sq8.mojo:567-582): the
        retained FP32 row i."""
        out = np.empty(self.d, np.float32)
        _lib.check(_lib.load().mvg_sq8_get_vector(self._h, int(i), _ptr(out)))
        return out

    def get_vectors(self, ids):
        """Retained FP32 rows ids[0..n) as [n, d]."""
        ids = np.ascontiguousarray(ids, np.int64).reshape(-1)
        out = np.empty((ids.size, self.d), np.float32)
        _lib.check(_lib.load().mvg_sq8_get_vectors(self._h, _ptr(ids), ids.size, _ptr(out)))
        return out

    def search(self, x, k: int, k_factor: int = 1, exclude=None):
        x = _f32(x, self.d)
        nq = x.shape[0]
        D = np.empty((nq, k), np.float32)
        I = np.empty((nq, k), np.int64)
        m, ml = _mask(exclude)
        _lib.check(_lib.load().mvg_sq8_search(self._h, _ptr(x), nq, k, k_factor, _ptr(m), ml,
                                               _ptr(D), _ptr(I)))
        return D, I


class KMeans:
    """Lloyd k-means (KMeans, kmeans.mojo:8-220) on the GPU."""

    def __init__(self, d: int, k: int, niter: int = 15, seed: int = KMEANS_SEED):
        self.d, self.k, self.niter, self.seed = d, k, niter, seed
        self.centroids = None
        self.assignments = None

    def train(self, x, init=None, sample_offset: int = 0):
        x = _f32(x, self.d)
        c = np.empty((self.k, self.d), np.float32)
        a = np.empty(x.shape[0], np.int32)
        ini = None if init is None else _f32(init, self.d)
        _lib.check(_lib.load().mvg_kmeans_train(_ptr(x), x.shape[0], self.d, self.k, self.niter,
                                                 _ptr(ini), sample_offset, self.seed, _ptr(c), _ptr(a)))
        self.centroids, self.assignments = c, a
        return c


class ProductQuantizer:
    """Product quantizer with 8-bit codes (ProductQuantizer, pq.mojo:80-292)."""

    def __init__(self, d: int, M: int, niter: int = 15, seed: int = KMEANS_SEED):
        if d <= 0:
            raise _lib.MojoVecError(1, "PQ dimension must be positive.")
        if M <= 0 or M > d or d % M != 0:
            raise _lib.MojoVecError(1, "PQ M must divide dimension and be between 1 and dimension.")
        self.d, self.M, self.ksub, self.dsub = d, M, 256, d // M
        self.niter, self.seed = niter, seed
        self.centroids = None
        self.is_trained = False

    def train(self, x):
        x = _f32(x, self.d)
        cb = np.empty((self.M, 256, self.dsub), np.float32)
        _lib.check(_lib.load().mvg_pq_train(_ptr(x), x.shape[0], self.d, self.M, self.niter,
                                             self.seed, _ptr(cb)))
        self.centroids, self.is_trained = cb, True

    def compute_codes(self, x):
        x = _f32(x, self.d)
        codes = np.empty((x.shape[0], self.M), np.uint8)
        cb = np.ascontiguousarray(self.centroids, np.float32)
        _lib.check(_lib.load().mvg_pq_encode(_ptr(x), x.shape[0], self.d, self.M, _ptr(cb), _ptr(codes)))
        return codes

    def decode(self, codes):
        """ProductQuantizer.decode (pq.mojo:221-250) on the device."""
        codes = np.ascontiguousarray(codes, np.uint8).reshape(-1, self.M)
        out = np.empty((codes.shape[0], self.d), np.float32)
        cb = np.ascontiguousarray(self.centroids, np.float32)
        _lib.check(_lib.load().mvg_pq_decode(_ptr(codes), codes.shape[0], self.d, self.M, _ptr(cb), _ptr(out)))
        return out

    def compute_distance_table(self, q, metric: int = METRIC_L2):
        q = _f32(q, self.d)
        lut = np.empty((q.shape[0], self.M, 256), np.float32)
        cb = np.ascontiguousarray(self.centroids, np.float32)
        _lib.check(_lib.load().mvg_pq_lut(_ptr(q), q.shape[0], self.d, self.M, _ptr(cb), metric, _ptr(lut)))
        return lut


class IndexIVFPQ:
    """IVF-PQ index with an HBM-resident CSR list layout (IndexIVFPQ,
    index_ivf_pq.mojo:12-360)."""

    def __init__(self, d: int, nlist: int, M: int, metric: int = METRIC_L2, keep_vectors: bool = True,
                 _handle=None):
        L = _lib.load()
        if _handle is None:
            h = ctypes.c_void_p()
            _lib.check(L.mvg_ivfpq_create(d, nlist, M, metric, ctypes.byref(h)))
            _lib.check(L.mvg_ivfpq_set_keep_vectors(h, int(keep_vectors)))
            self._owned = True
        else:
            h = ctypes.c_void_p(_handle)
            self._owned = False
        self._h, self.d, self.nlist, self.M, self.metric_type = h, d, nlist, M, metric

    def __del__(self):
        if getattr(self, "_h", None) and getattr(self, "_owned", False):
            _lib.load().mvg_ivfpq_destroy(self._h)
            self._h = None

    def _info(self):
        n, t, p = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.load().mvg_ivfpq_info(self._h, ctypes.byref(n), ctypes.byref(t), ctypes.byref(p)))
        return n.value, bool(t.value), p.value

    @property
    def ntotal(self) -> int:
        return self._info()[0]

    @property
    def is_trained(self) -> bool:
        return self._info()[1]

    @property
    def nprobe(self) -> int:
        return self._info()[2]

    @nprobe.setter
    def nprobe(self, v: int) -> None:
        _lib.check(_lib.load().mvg_ivfpq_set_nprobe(self._h, int(v)))

    def set_search_mode(self, precision: str = "plain", split_scan: bool = False) -> None:
        """precision "plain" (reference FP32 order) or "mixed" (decomposed
        tables; pair with k_factor > 1 for the exact rerank)."""
        p = {"plain": 0, "mixed": 1}[precision]
        _lib.check(_lib.load().mvg_ivfpq_set_search_mode(self._h, p, int(split_scan)))

    def set_prefilter(self, per_k: int = 8) -> None:
        """Mixed form: narrow the k * k_factor candidates to max(64, per_k * k)
        by the SQ8 code distance of the stored rows before the exact FP32
        rerank; 0 reranks every candidate in FP32."""
        _lib.check(_lib.load().mvg_ivfpq_set_prefilter(self._h, int(per_k)))

    def set_table_bits(self, bits: int = 32) -> None:
        """Mixed form, L2 scans that the exact rerank follows (k_factor > 1):
        8 sums 8-bit tables in the deep-pool scan; 16 runs the FP16 MFMA tile
        scan (query tiles times PQ-decoded entries; PQ with 8-dimensional
        subspaces, M in {8, 12, 16}) on batches of at least the tile minimum
        and the 8-bit scan below it; 32 keeps FP32 tables."""
        _lib.check(_lib.load().mvg_ivfpq_set_table_bits(self._h, int(bits)))

    def set_tile_min_batch(self, min_queries: int = 0) -> None:
        """Fewest queries of a batch for the FP16 tile scan (table bits 16);
        0: 16 per compute unit."""
        _lib.check(_lib.load().mvg_ivfpq_set_tile_min_batch(self._h, int(min_queries)))

    def set_work_queue(self, entries_per_item: int = 0) -> None:
        """Persistent (query, entry-range) work queue for skewed lists; 0 off."""
        _lib.check(_lib.load().mvg_ivfpq_set_work_queue(self._h, int(entries_per_item)))

    def set_deep_select(self, min_candidates: int = 0, pool: str = "hbm") -> None:
        """k * k_factor from which the scan selects candidates without a
        sorted running top set (0: default 513); pool "hbm" stores every key
        in HBM and selects once per query, "lds" re-thresholds a running set
        in LDS. Same results."""
        if pool not in ("lds", "hbm"):
            raise ValueError('pool must be "lds" or "hbm"')
        _lib.check(_lib.load().mvg_ivfpq_set_deep_select(self._h, int(min_candidates),
                                                         0 if pool == "hbm" else 1))

    def train(self, x, niter_coarse: int = 15, niter_pq: int = 15, seed: int = KMEANS_SEED) -> None:
        x = _f32(x, self.d)
        _lib.check(_lib.load().mvg_ivfpq_train(self._h, _ptr(x), x.shape[0], niter_coarse, niter_pq, seed))

    def set_codebooks(self, coarse, pq) -> None:
        c = _f32(coarse, self.d)
        p = np.ascontiguousarray(pq, np.float32)
        _lib.check(_lib.load().mvg_ivfpq_set_codebooks(self._h, _ptr(c), _ptr(p)))

    def codebooks(self):
        c = np.empty((self.nlist, self.d), np.float32)
        p = np.empty((self.M, 256, self.d // self.M), np.float32)
        _lib.check(_lib.load().mvg_ivfpq_get_codebooks(self._h, _ptr(c), _ptr(p)))
        return c, p

    def add(self, x) -> None:
        x = _f32(x, self.d)
        _lib.check(_lib.load().mvg_ivfpq_add(self._h, _ptr(x), x.shape[0]))

    def add_with_ids(self, x, ids) -> None:
        """IndexIVFPQ.add_with_ids (index_ivf_pq.mojo:150-213): ids in
        [0, 2^31 - 1], unique while FP32 rows are kept."""
        x = _f32(x, self.d)
        ids = np.ascontiguousarray(ids, np.int64).reshape(-1)
        if ids.size != x.shape[0]:
            raise ValueError("add_with_ids needs one id per vector")
        _lib.check(_lib.load().mvg_ivfpq_add_with_ids(self._h, _ptr(x), x.shape[0], _ptr(ids)))

    def save(self, path) -> None:
        """write_index_ivf_pq (serialization.mojo:1029-1044) into a file."""
        _lib.check(_lib.load().mvg_ivfpq_save(self._h, os.fspath(path).encode()))

    @classmethod
    def load(cls, path) -> "IndexIVFPQ":
        """read_index_ivf_pq (serialization.mojo:1046-1090) from a file; the
        lists and codebooks go to HBM. Snapshots hold no FP32 rows, so a loaded
        index searches with k_factor 1."""
        L = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(L.mvg_ivfpq_load(os.fspath(path).encode(), ctypes.byref(h)))
        vals = [ctypes.c_int() for _ in range(4)]
        _lib.check(L.mvg_ivfpq_shape(h, *[ctypes.byref(v) for v in vals]))
        ix = cls.__new__(cls)
        ix._h, ix._owned = h, True
        ix.d, ix.nlist, ix.M, ix.metric_type = (v.value for v in vals)
        return ix

    def list_sizes(self):
        s = np.empty(self.nlist, np.int64)
        _lib.check(_lib.load().mvg_ivfpq_list_sizes(self._h, _ptr(s)))
        return s

    def get_list(self, l: int):
        n = int(self.list_sizes()[l])
        ids = np.empty(n, np.int64)
        codes = np.empty((n, self.M), np.uint8)
        _lib.check(_lib.load().mvg_ivfpq_get_list(self._h, l, _ptr(ids), _ptr(codes)))
        return ids, codes

    def search(self, x, k: int, k_factor: int = 1, nprobe_per_query=None, exclude=None):
        x = _f32(x, self.d)
        nq = x.shape[0]
        D = np.empty((nq, k), np.float32)
        I = np.empty((nq, k), np.int64)
        npq = None if nprobe_per_query is None else np.ascontiguousarray(nprobe_per_query, np.int32)
        m, ml = _mask(exclude)
        _lib.check(_lib.load().mvg_ivfpq_search(self._h, _ptr(x), nq, k, _ptr(npq), k_factor,
                                                 _ptr(m), ml, _ptr(D), _ptr(I)))
        return D, I

    SCHEDULES = ("per_query", "queue_1024", "queue_4096", "queue_16384", "per_list")

    def autotune(self, q_ptr: int, nq: int, k: int, D_ptr: int, I_ptr: int, k_factor: int = 1,
                 nprobe_ptr: int = 0, max_nprobe: int = 0):
        """Time the scan schedules on this batch (device pointers), keep the
        fastest; returns (schedule name, ms per search)."""
        c, ms = ctypes.c_int(), ctypes.c_float()
        _lib.check(_lib.load().mvg_ivfpq_autotune(
            self._h, ctypes.c_void_p(q_ptr), nq, k, ctypes.c_void_p(nprobe_ptr or None), max_nprobe,
            k_factor, ctypes.c_void_p(D_ptr), ctypes.c_void_p(I_ptr), ctypes.byref(c), ctypes.byref(ms)))
        return self.SCHEDULES[c.value], ms.value

    def capture(self, q_ptr: int, nq: int, k: int, D_ptr: int, I_ptr: int, k_factor: int = 1,
                nprobe_ptr: int = 0, max_nprobe: int = 0, kq_ptr: int = 0) -> "SearchGraph":
        """Record one device-pointer search of this batch shape as a HIP graph.
        kq_ptr (device int32 [nq]): a k per query, k being the largest (the
        output row stride)."""
        g = ctypes.c_void_p()
        _lib.check(_lib.load().mvg_ivfpq_capture_k(
            self._h, ctypes.c_void_p(q_ptr), nq, ctypes.c_void_p(kq_ptr or None), k,
            ctypes.c_void_p(nprobe_ptr or None), max_nprobe, k_factor, ctypes.c_void_p(D_ptr),
            ctypes.c_void_p(I_ptr), ctypes.byref(g)))
        return SearchGraph(g, self)

    def search_device(self, q_ptr: int, nq: int, k: int, D_ptr: int, I_ptr: int, k_factor: int = 1,
                      nprobe_ptr: int = 0, max_nprobe: int = 0, stream: int = 0, kq_ptr: int = 0) -> None:
        """Search with every buffer already in device memory (raw pointers);
        kq_ptr as in capture."""
        _lib.check(_lib.load().mvg_ivfpq_search_device_k(
            self._h, ctypes.c_void_p(q_ptr), nq, ctypes.c_void_p(kq_ptr or None), k,
            ctypes.c_void_p(nprobe_ptr or None), max_nprobe, k_factor, None, 0, ctypes.c_void_p(D_ptr),
            ctypes.c_void_p(I_ptr), ctypes.c_void_p(stream or None)))


class IndexIVFFlat:
    """IVF index over exact FP32 rows with an HBM-resident CSR list layout
    (IndexIVFFlat, index_ivf_flat.mojo:10-244)."""

    def __init__(self, d: int, nlist: int, metric: int = METRIC_L2, quantizer_metric=None):
        """quantizer_metric: the coarse IndexFlat's metric (default: metric)."""
        qm = metric if quantizer_metric is None else quantizer_metric
        h = ctypes.c_void_p()
        _lib.check(_lib.load().mvg_ivfflat_create_with_quantizer(d, nlist, metric, qm, ctypes.byref(h)))
        self._h, self.d, self.nlist, self.metric_type, self.quantizer_metric = h, d, nlist, metric, qm

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.load().mvg_ivfflat_destroy(self._h)
            self._h = None

    def _info(self):
        n, t, p = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.load().mvg_ivfflat_info(self._h, ctypes.byref(n), ctypes.byref(t), ctypes.byref(p)))
        return n.value, bool(t.value), p.value

    @property
    def ntotal(self) -> int:
        return self._info()[0]

    @property
    def is_trained(self) -> bool:
        return self._info()[1]

    @property
    def nprobe(self) -> int:
        return self._info()[2]

    @nprobe.setter
    def nprobe(self, v: int) -> None:
        _lib.check(_lib.load().mvg_ivfflat_set_nprobe(self._h, int(v)))

    def train(self, x, niter: int = 15, seed: int = KMEANS_SEED) -> None:
        """KMeans(d, nlist, 15).train over x (index_ivf_flat.mojo:65-77)."""
        x = _f32(x, self.d)
        _lib.check(_lib.load().mvg_ivfflat_train(self._h, _ptr(x), x.shape[0], niter, seed))

    def set_centroids(self, coarse) -> None:
        c = _f32(coarse, self.d)
        _lib.check(_lib.load().mvg_ivfflat_set_centroids(self._h, _ptr(c)))

    def centroids(self):
        c = np.empty((self.nlist, self.d), np.float32)
        _lib.check(_lib.load().mvg_ivfflat_get_centroids(self._h, _ptr(c)))
        return c

    def add(self, x) -> None:
        x = _f32(x, self.d)
        _lib.check(_lib.load().mvg_ivfflat_add(self._h, _ptr(x), x.shape[0], None))

    def add_with_ids(self, x, ids) -> None:
        x = _f32(x, self.d)
        ids = np.ascontiguousarray(ids, np.int64)
        if ids.shape != (x.shape[0],):
            raise ValueError("one id per vector is required")
        _lib.check(_lib.load().mvg_ivfflat_add(self._h, _ptr(x), x.shape[0], _ptr(ids)))

    def save(self, path) -> None:
        """write_index_ivf_flat (serialization.mojo:980-993) into a file."""
        _lib.check(_lib.load().mvg_ivfflat_save(self._h, os.fspath(path).encode()))

    @classmethod
    def load(cls, path) -> "IndexIVFFlat":
        """read_index_ivf_flat (serialization.mojo:995-1020) from a file."""
        L = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(L.mvg_ivfflat_load(os.fspath(path).encode(), ctypes.byref(h)))
        vals = [ctypes.c_int() for _ in range(4)]
        _lib.check(L.mvg_ivfflat_shape(h, *[ctypes.byref(v) for v in vals]))
        ix = cls.__new__(cls)
        ix._h = h
        ix.d, ix.nlist, ix.metric_type, ix.quantizer_metric = (v.value for v in vals)
        return ix

    def list_sizes(self):
        s = np.empty(self.nlist, np.int64)
        _lib.check(_lib.load().mvg_ivfflat_list_sizes(self._h, _ptr(s)))
        return s

    def get_list(self, l: int):
        n = int(self.list_sizes()[l])
        ids = np.empty(n, np.int64)
        rows = np.empty((n, self.d), np.float32)
        _lib.check(_lib.load().mvg_ivfflat_get_list(self._h, l, _ptr(ids), _ptr(rows)))
        return ids, rows

    def search(self, x, k: int, exclude=None):
        x = _f32(x, self.d)
        nq = x.shape[0]
        D = np.empty((nq, k), np.float32)
        I = np.empty((nq, k), np.int64)
        m, ml = _mask(exclude)
        _lib.check(_lib.load().mvg_ivfflat_search(self._h, _ptr(x), nq, k, _ptr(m), ml, _ptr(D),
                                                   _ptr(I)))
        return D, I

    def search_device(self, q_ptr: int, nq: int, k: int, D_ptr: int, I_ptr: int,
                      stream: int = 0) -> None:
        """Search with every buffer already in device memory (raw pointers)."""
        _lib.check(_lib.load().mvg_ivfflat_search_device(
            self._h, ctypes.c_void_p(q_ptr), nq, k, None, 0, ctypes.c_void_p(D_ptr),
            ctypes.c_void_p(I_ptr), ctypes.c_void_p(stream or None)))


class SearchGraph:
    """A captured IVF-PQ search batch (mvg_ivfpq_capture); replay with launch()."""

    def __init__(self, handle, index):
        self._h, self._index = handle, index

    def launch(self, stream: int = 0) -> None:
        _lib.check(_lib.load().mvg_graph_launch(self._h, ctypes.c_void_p(stream or None)))

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.load().mvg_graph_destroy(self._h)
            self._h = None
