"""IVFPQCollection: drop-in mirror of MojoVec's Python facade
(python/mojovec/_ivfpq.py:28-218) over the GPU collection
(mojovec_gpu.h mvg_coll_*, restating mojovec/api/collection_ivfpq.mojo).

Same constructor, methods, defaults, result dictionaries and error messages;
`k_factor` (default 1 = reference behaviour) adds the exact FP32 rerank.
"""
from __future__ import annotations

import ctypes
import os
from collections.abc import Sequence
from numbers import Real
from typing import Any, TypedDict

import numpy as np

from . import _lib

_METRIC_NAMES = {0: "l2", 1: "ip", 2: "cosine"}


class IVFPQStats(TypedDict):
    count: int
    dimension: int
    nlist: int
    M: int
    nprobe: int
    trained: bool
    metric: str
    code_size_bytes: int


class QueryResult(TypedDict):
    ids: list
    distances: list
    metadatas: list
    documents: list
    scores: list


def _flatten_embeddings(embeddings, dimension=None):
    """Restates python/mojovec/_collection.py:21-52."""
    if isinstance(embeddings, np.ndarray):
        embeddings = embeddings.tolist()
    if len(embeddings) == 0:
        return []
    first = embeddings[0]
    if isinstance(first, Real):
        if dimension is not None and len(embeddings) % dimension != 0:
            raise ValueError(
                f"flat embeddings contain {len(embeddings)} values; "
                f"expected a multiple of dimension {dimension}"
            )
        return embeddings
    flat: list[float] = []
    for index, row in enumerate(embeddings):
        if isinstance(row, Real):
            raise ValueError("embeddings must be homogenously flat or 2D")
        row_len = len(row)
        if dimension is not None and row_len != dimension:
            raise ValueError(
                f"embedding at index {index} has dimension {row_len}; expected {dimension}"
            )
        flat.extend(float(v) for v in row)
    return flat


def _validated_numpy_vectors(ids, embeddings, dimension):
    """Restates python/mojovec/_collection.py:66-92."""
    ids_array = np.asarray(ids)
    emb = np.asarray(embeddings)
    if ids_array.dtype != np.int64 or ids_array.ndim != 1 or not ids_array.flags.c_contiguous:
        raise TypeError("ids must be a contiguous one-dimensional numpy.int64 array")
    if emb.dtype != np.float32 or emb.ndim not in (1, 2) or not emb.flags.c_contiguous:
        raise TypeError("embeddings must be a contiguous numpy.float32 array")
    if emb.size != ids_array.size * dimension:
        raise ValueError("embeddings shape does not match ids and dimension")
    return ids_array, emb


class IVFPQCollection:
    """Compressed IVF-PQ vector collection resident in GPU memory.

    Parameters match python/mojovec/_ivfpq.py:53-70: dimension, nlist=100,
    pq_subvectors=16, nprobe=None (min(10, nlist)), metric="l2", name="".
    """

    def __init__(self, dimension: int, nlist: int = 100, pq_subvectors: int = 16,
                 nprobe: int | None = None, metric: str = "l2", name: str = "",
                 k_factor: int = 1) -> None:
        L = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(L.mvg_coll_create(int(dimension), int(nlist), int(pq_subvectors),
                                     0 if nprobe is None else int(nprobe), name.encode(),
                                     metric.encode(), ctypes.byref(h)))
        self._h = h
        self._k_factor = int(k_factor)

    @classmethod
    def load(cls, path) -> "IVFPQCollection":
        """Load an owned IVF-PQ snapshot (collection_ivfpq.mojo:372-476)."""
        h = ctypes.c_void_p()
        _lib.check(_lib.load().mvg_coll_load(os.fspath(path).encode(), ctypes.byref(h)))
        collection = cls.__new__(cls)
        collection._h = h
        collection._k_factor = 1
        return collection

    def save(self, path) -> None:
        """Atomically persist a checksummed owned IVF-PQ snapshot in the
        reference's format (collection_ivfpq.mojo:335-370)."""
        _lib.check(_lib.load().mvg_coll_save(self._h, os.fspath(path).encode()))

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.load().mvg_coll_destroy(self._h)
            self._h = None

    def _info(self):
        vals = [ctypes.c_int(), ctypes.c_int(), ctypes.c_int64(), ctypes.c_int(), ctypes.c_int(),
                ctypes.c_int(), ctypes.c_int()]
        _lib.check(_lib.load().mvg_coll_info(self._h, *[ctypes.byref(v) for v in vals]))
        return [v.value for v in vals]

    def __repr__(self) -> str:
        return (
            f"IVFPQCollection(name={self.name()!r}, "
            f"dimension={self.dimension()}, metric={self.metric()!r}, "
            f"count={self.count()}, nlist={self.nlist()}, "
            f"pq_subvectors={self.pq_subvectors()}, "
            f"nprobe={self.nprobe()}, trained={self.is_trained()})"
        )

    def name(self) -> str:
        buf = ctypes.create_string_buffer(4096)
        _lib.check(_lib.load().mvg_coll_name(self._h, buf, len(buf)))
        return buf.value.decode()

    def dimension(self) -> int:
        return self._info()[0]

    def metric(self) -> str:
        return _METRIC_NAMES[self._info()[1]]

    def count(self) -> int:
        return self._info()[2]

    def is_trained(self) -> bool:
        return bool(self._info()[6])

    def nlist(self) -> int:
        return self._info()[3]

    def pq_subvectors(self) -> int:
        return self._info()[4]

    def nprobe(self) -> int:
        return self._info()[5]

    def set_nprobe(self, nprobe: int) -> None:
        _lib.check(_lib.load().mvg_coll_set_nprobe(self._h, int(nprobe)))

    def set_k_factor(self, k_factor: int) -> None:
        if k_factor < 1:
            raise ValueError("k_factor must be at least 1")
        self._k_factor = int(k_factor)

    def set_train_params(self, niter_coarse: int = 15, niter_pq: int = 15,
                         seed: int = 0x4D4F4A4F564543) -> None:
        _lib.check(_lib.load().mvg_coll_set_train_params(self._h, niter_coarse, niter_pq, seed))

    def stats(self) -> IVFPQStats:
        d, m, c, nl, M, np_, t = self._info()
        return {"count": c, "dimension": d, "nlist": nl, "M": M, "nprobe": np_,
                "trained": bool(t), "metric": _METRIC_NAMES[m], "code_size_bytes": M}

    def index(self):
        """The underlying IndexIVFPQ (borrowed handle)."""
        from .index import IndexIVFPQ
        d, m, _, nl, M, _, _ = self._info()
        return IndexIVFPQ(d, nl, M, 0 if m == 0 else 1,
                          _handle=_lib.load().mvg_coll_index(self._h))

    def _train_flat(self, flat: np.ndarray) -> None:
        _lib.check(_lib.load().mvg_coll_train(self._h, ctypes.c_void_p(flat.ctypes.data), flat.size))

    def train(self, embeddings: Sequence[Real] | Sequence[Sequence[Real]]) -> None:
        flat = np.asarray(_flatten_embeddings(embeddings, self.dimension()), np.float32)
        self._train_flat(flat)

    def train_numpy(self, embeddings: Any) -> None:
        array = np.asarray(embeddings)
        if (array.dtype != np.float32 or array.ndim != 2 or array.shape[1] != self.dimension()
                or not array.flags.c_contiguous):
            raise TypeError("embeddings must be a contiguous (n, dimension) numpy.float32 array")
        self._train_flat(array)

    def _add_flat(self, ids: np.ndarray, flat: np.ndarray) -> None:
        _lib.check(_lib.load().mvg_coll_add(self._h, ctypes.c_void_p(ids.ctypes.data), ids.size,
                                            ctypes.c_void_p(flat.ctypes.data), flat.size))

    def add(self, ids: Sequence[int], embeddings) -> None:
        flat = _flatten_embeddings(embeddings, self.dimension())
        expected = len(ids) * self.dimension()
        if len(flat) != expected:
            raise ValueError(f"embeddings contain {len(flat)} values; expected {expected}")
        self._add_flat(np.asarray(list(ids), np.int64), np.asarray(flat, np.float32))

    def add_numpy(self, ids: Any, embeddings: Any) -> None:
        ids_array, emb = _validated_numpy_vectors(ids, embeddings, self.dimension())
        self._add_flat(ids_array, emb.reshape(-1))

    def _query_flat(self, flat: np.ndarray, n_results: int, k_factor: int | None) -> QueryResult:
        d = self.dimension()
        nq = flat.size // d if flat.size % d == 0 else 0
        ids = np.empty(max(nq * n_results, 1), np.int64)
        dist = np.empty(max(nq * n_results, 1), np.float32)
        nq_out = ctypes.c_int64()
        kf = self._k_factor if k_factor is None else int(k_factor)
        _lib.check(_lib.load().mvg_coll_query(self._h, ctypes.c_void_p(flat.ctypes.data), flat.size,
                                              int(n_results), kf, ctypes.c_void_p(ids.ctypes.data),
                                              ctypes.c_void_p(dist.ctypes.data), ctypes.byref(nq_out)))
        nq = nq_out.value
        ids = ids[: nq * n_results].reshape(nq, n_results)
        dist = dist[: nq * n_results].reshape(nq, n_results)
        return {"ids": ids.tolist(), "distances": dist.tolist(), "metadatas": [], "documents": [],
                "scores": []}

    def query(self, query_embeddings, n_results: int = 10, k_factor: int | None = None) -> QueryResult:
        flat = np.asarray(_flatten_embeddings(query_embeddings, self.dimension()), np.float32)
        return self._query_flat(flat, n_results, k_factor)

    def query_numpy(self, query_embeddings: Any, n_results: int = 10,
                    k_factor: int | None = None) -> QueryResult:
        array = np.asarray(query_embeddings)
        if array.dtype != np.float32:
            raise TypeError("query_embeddings must have numpy.float32 dtype")
        if array.ndim == 1:
            if array.size % self.dimension() != 0:
                raise ValueError("query embeddings length must be a multiple of dimension")
            array = array.reshape((-1, self.dimension()))
        if array.ndim != 2 or array.shape[1] != self.dimension() or not array.flags.c_contiguous:
            raise ValueError("query_embeddings must be a contiguous (n, dimension) array")
        return self._query_flat(array.reshape(-1), n_results, k_factor)
