"""Continuous-batching server over HBM-resident IVF-PQ and Flat SQ8 indexes
(mvg_server_* in include/mojovec_gpu.h).

    srv = Server(index, slots=4, max_batch=8192, mem_budget=8 << 30)
    t = srv.submit(queries, k=10, nprobe=32, k_factor=4)
    D, I, latency_ms = srv.wait(t)
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


class Server:
    """Continuous batching over one index, or over an IVF-PQ and a Flat SQ8
    index together (Server.both): one dispatcher, one set of slots and one
    memory budget; submit(..., target=index) names the index of a request."""

    def __init__(self, index, slots: int = 4, max_batch: int = 8192, mem_budget: int = 8 << 30,
                 _sq8=None):
        from .index import IndexFlatSQ8

        h = ctypes.c_void_p()
        if isinstance(index, IndexFlatSQ8):
            ivf, sq = None, index
        else:
            ivf, sq = index, _sq8
        _lib.check(_lib.load().mvg_server_create_multi(
            None if ivf is None else ivf._h, None if sq is None else sq._h, slots, max_batch, mem_budget,
            ctypes.byref(h)))
        self._h = h
        self._ivf, self._sq8 = ivf, sq  # keeps the indexes alive
        self._index = ivf if ivf is not None else sq
        self._out = {}

    @classmethod
    def both(cls, ivfpq, sq8, slots: int = 4, max_batch: int = 8192, mem_budget: int = 8 << 30):
        return cls(ivfpq, slots, max_batch, mem_budget, _sq8=sq8)

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.load().mvg_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def submit(self, queries, k: int = 10, nprobe: int = 8, k_factor: int = 1, target=None) -> int:
        if target is None:
            target = self._index
        t_idx = 1 if target is self._sq8 and target is not None else 0
        q = np.ascontiguousarray(queries, dtype=np.float32).reshape(-1, target.d)
        nq = q.shape[0]
        D = np.empty((nq, k), np.float32)
        I = np.empty((nq, k), np.int64)
        t = ctypes.c_int64()
        _lib.check(_lib.load().mvg_server_submit_to(self._h, t_idx, ctypes.c_void_p(q.ctypes.data), nq, k,
                                                    nprobe, k_factor, ctypes.c_void_p(D.ctypes.data),
                                                    ctypes.c_void_p(I.ctypes.data), ctypes.byref(t)))
        self._out[t.value] = (D, I)
        return t.value

    def wait(self, ticket: int):
        lat = ctypes.c_double()
        rc = _lib.load().mvg_server_wait(self._h, ticket, ctypes.byref(lat))
        D, I = self._out.pop(ticket)
        _lib.check(rc)
        return D, I, lat.value

    def stats(self):
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.check(_lib.load().mvg_server_stats(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"completed_queries": a.value, "batches": b.value, "max_inflight_batches": c.value}
