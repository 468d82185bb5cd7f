"""Restatement of MojoVec's IVF-PQ and IVF-Flat snapshot file formats, for
tests only.

TEST INFRASTRUCTURE. Only tests/ may import this module; the product reads
and writes snapshots in mojovec_amd/csrc/snapshot.cpp.

Layout (all Int fields 8-byte little-endian, Bool one byte, floats raw FP32):

  collection  mojovec/api/collection_ivfpq.mojo:349-362
      magic 0x43495650, version 1, dimension, name (len + UTF-8),
      metric name (len + UTF-8), num_ids, user ids [num_ids] Int64,
      IndexIVFPQ stream, then the checksum trailer
  IndexIVFPQ  mojovec/io/serialization.mojo:1029-1044
      magic 0x4d4a4f50, d, nlist, M, nprobe, ntotal, is_trained (Bool),
      metric (0 L2 / 1 IP), IndexFlat, ArrayInvertedLists, ProductQuantizer
  IndexIVFFlat  serialization.mojo:980-993
      magic 0x4d4a4f49, d, nlist, nprobe, ntotal, is_trained (Bool), metric,
      IndexFlat (with the quantizer's own metric), ArrayInvertedLists with
      code_size 4*d whose codes are the raw FP32 rows
  IndexFlat   serialization.mojo:229-238
      magic 0x4d4a4f46, d, ntotal, capacity, metric, capacity*d floats
  invlists    serialization.mojo:903-914
      magic 0x4d4a4f4c, nlist, code_size, per list: size, capacity,
      ids [size] Int64, codes [size*code_size]
  PQ          serialization.mojo:946-952
      magic 0x4d4a4f51, d, M, ksub, is_trained (Bool), M*ksub*dsub floats
  trailer     mojovec/io/snapshot_file.mojo:74-84
      0x4D4F4A4F534E4150, checksum of the payload (_checksum_prefix :45-59,
      8 MiB chunks, _checksum_update :19-42)

Capacities follow the reference's storage: the coarse IndexFlat starts at
1024 rows and grows to max(2*cap, n) (index_flat.mojo:134,199-206); a list
grows 0 -> max(16, n) -> doubling, one append at a time
(inverted_lists.mojo:119-150). Rows past the coarse ntotal and untrained PQ
centroids are unset memory in the reference; they are written as zeros here.
"""
from __future__ import annotations

import struct

import numpy as np

MAGIC_FLAT = 0x4D4A4F46
MAGIC_IVF_FLAT = 0x4D4A4F49
MAGIC_IVF_PQ = 0x4D4A4F50
MAGIC_INVLISTS = 0x4D4A4F4C
MAGIC_PQ = 0x4D4A4F51
COLL_MAGIC = 0x43495650
COLL_VERSION = 1
SUM_MAGIC = 0x4D4F4A4F534E4150
SUM_SEED = 0xD6E8FEB86659FD93
SUM_CHUNK = 8 * 1024 * 1024
_M64 = (1 << 64) - 1


def _rotl(v: int, a: int) -> int:
    return ((v << a) | (v >> (64 - a))) & _M64


def checksum_chunk(state: int, chunk: bytes) -> int:
    """_checksum_update (snapshot_file.mojo:19-42)."""
    n = len(chunk)
    words = n // 8
    for w in np.frombuffer(chunk[: words * 8], dtype="<u8").tolist():
        state ^= w
        state = _rotl(state, 27) ^ _rotl(w, 17)
        state ^= state >> 23
    tail = int.from_bytes(chunk[words * 8:], "little")
    state ^= tail
    state ^= (n << 32) & _M64
    return _rotl(state, 11)


def checksum(payload: bytes) -> int:
    """_checksum_prefix (snapshot_file.mojo:45-59) over a whole payload."""
    state = SUM_SEED
    for off in range(0, len(payload), SUM_CHUNK):
        state = checksum_chunk(state, payload[off: off + SUM_CHUNK])
    return state


def list_capacity(size: int, cap: int = 0) -> int:
    """Reference list capacity after appending one entry at a time from
    capacity `cap` up to `size` entries (inverted_lists.mojo:119-130)."""
    while cap < size:
        cap = 16 if cap == 0 else 2 * cap
    return cap


def flat_capacity(rows: int, cap: int = 1024) -> int:
    return max(2 * cap, rows) if rows > cap else cap


def _i64(v: int) -> bytes:
    return struct.pack("<q", int(v))


def index_bytes(d, nlist, M, nprobe, ntotal, trained, metric, coarse, lists, pq,
                coarse_capacity=None, list_caps=None) -> bytes:
    """write_index_ivf_pq. `lists` is a sequence of (ids int64[n], codes
    uint8[n, M]); coarse [nlist, d] (ignored when untrained); pq [M, 256, d/M]."""
    out = bytearray()
    out += _i64(MAGIC_IVF_PQ) + _i64(d) + _i64(nlist) + _i64(M) + _i64(nprobe) + _i64(ntotal)
    out += bytes([1 if trained else 0]) + _i64(metric)
    rows = nlist if trained else 0
    cap = flat_capacity(rows) if coarse_capacity is None else coarse_capacity
    out += _i64(MAGIC_FLAT) + _i64(d) + _i64(rows) + _i64(cap) + _i64(metric)
    cb = np.zeros((cap, d), np.float32)
    if rows:
        cb[:rows] = np.asarray(coarse, np.float32).reshape(rows, d)
    out += cb.tobytes()
    out += _i64(MAGIC_INVLISTS) + _i64(nlist) + _i64(M)
    for l, (ids, codes) in enumerate(lists):
        ids = np.asarray(ids, np.int64)
        codes = np.asarray(codes, np.uint8).reshape(ids.size, M)
        c = list_capacity(ids.size) if list_caps is None else list_caps[l]
        out += _i64(ids.size) + _i64(c) + ids.astype("<i8").tobytes() + codes.tobytes()
    out += _i64(MAGIC_PQ) + _i64(d) + _i64(M) + _i64(256) + bytes([1 if trained else 0])
    p = np.asarray(pq, np.float32).reshape(-1) if trained else np.zeros(d * 256, np.float32)
    out += p.astype("<f4").tobytes()
    return bytes(out)


def ivfflat_bytes(d, nlist, nprobe, ntotal, trained, metric, coarse, lists, quantizer_metric=None,
                  coarse_capacity=None, list_caps=None) -> bytes:
    """write_index_ivf_flat. `lists` is a sequence of (ids int64[n], rows
    float32[n, d]); coarse [nlist, d] (ignored when untrained)."""
    qm = metric if quantizer_metric is None else quantizer_metric
    out = bytearray()
    out += _i64(MAGIC_IVF_FLAT) + _i64(d) + _i64(nlist) + _i64(nprobe) + _i64(ntotal)
    out += bytes([1 if trained else 0]) + _i64(metric)
    rows = nlist if trained else 0
    cap = flat_capacity(rows) if coarse_capacity is None else coarse_capacity
    out += _i64(MAGIC_FLAT) + _i64(d) + _i64(rows) + _i64(cap) + _i64(qm)
    cb = np.zeros((cap, d), np.float32)
    if rows:
        cb[:rows] = np.asarray(coarse, np.float32).reshape(rows, d)
    out += cb.tobytes()
    out += _i64(MAGIC_INVLISTS) + _i64(nlist) + _i64(4 * d)
    for l, (ids, vecs) in enumerate(lists):
        ids = np.asarray(ids, np.int64)
        vecs = np.asarray(vecs, np.float32).reshape(ids.size, d)
        c = list_capacity(ids.size) if list_caps is None else list_caps[l]
        out += _i64(ids.size) + _i64(c) + ids.astype("<i8").tobytes() + vecs.astype("<f4").tobytes()
    return bytes(out)


def parse_ivfflat(buf: bytes) -> dict:
    """Parses an IndexIVFFlat file into plain arrays."""
    cur = _Cursor(buf)
    assert cur.i64() == MAGIC_IVF_FLAT
    r = {k: cur.i64() for k in ("d", "nlist", "nprobe", "ntotal")}
    r["trained"] = cur.byte()
    r["metric"] = cur.i64()
    d = r["d"]
    assert cur.i64() == MAGIC_FLAT
    r["q_d"], r["q_rows"], r["q_capacity"], r["q_metric"] = (cur.i64() for _ in range(4))
    coarse = np.frombuffer(cur.take(4 * r["q_capacity"] * d), "<f4").reshape(r["q_capacity"], d)
    r["coarse"], r["coarse_tail"] = coarse[: r["q_rows"]], coarse[r["q_rows"]:]
    assert cur.i64() == MAGIC_INVLISTS
    r["l_nlist"], r["code_size"] = cur.i64(), cur.i64()
    lists, caps = [], []
    for _ in range(r["l_nlist"]):
        size, cap = cur.i64(), cur.i64()
        ids = np.frombuffer(cur.take(8 * size), "<i8")
        vecs = np.frombuffer(cur.take(size * r["code_size"]), "<f4").reshape(size, d)
        lists.append((ids, vecs))
        caps.append(cap)
    r["lists"], r["list_caps"] = lists, caps
    assert cur.pos == len(buf)
    return r


def collection_bytes(dimension, name, metric_name, user_ids, index: bytes) -> bytes:
    """CollectionIVFPQ.save payload plus its checksum trailer."""
    nb, mb = name.encode(), metric_name.encode()
    payload = bytearray()
    payload += _i64(COLL_MAGIC) + _i64(COLL_VERSION) + _i64(dimension)
    payload += _i64(len(nb)) + nb + _i64(len(mb)) + mb
    ids = np.asarray(user_ids, np.int64)
    payload += _i64(ids.size) + ids.astype("<i8").tobytes()
    payload += index
    return bytes(payload) + struct.pack("<QQ", SUM_MAGIC, checksum(bytes(payload)))


class _Cursor:
    def __init__(self, buf: bytes, pos: int = 0):
        self.buf, self.pos = buf, pos

    def i64(self) -> int:
        v = struct.unpack_from("<q", self.buf, self.pos)[0]
        self.pos += 8
        return v

    def byte(self) -> int:
        v = self.buf[self.pos]
        self.pos += 1
        return v

    def take(self, n: int) -> bytes:
        v = self.buf[self.pos: self.pos + n]
        assert len(v) == n, "truncated"
        self.pos += n
        return v


def parse_index(cur: _Cursor) -> dict:
    """Parses an IndexIVFPQ stream into plain arrays (no validation beyond
    magics: it reads what the product wrote)."""
    assert cur.i64() == MAGIC_IVF_PQ
    r = {k: cur.i64() for k in ("d", "nlist", "M", "nprobe", "ntotal")}
    r["trained"] = cur.byte()
    r["metric"] = cur.i64()
    d = r["d"]
    assert cur.i64() == MAGIC_FLAT
    r["q_d"], r["q_rows"], r["q_capacity"], r["q_metric"] = (cur.i64() for _ in range(4))
    coarse = np.frombuffer(cur.take(4 * r["q_capacity"] * d), "<f4").reshape(r["q_capacity"], d)
    r["coarse"], r["coarse_tail"] = coarse[: r["q_rows"]], coarse[r["q_rows"]:]
    assert cur.i64() == MAGIC_INVLISTS
    r["l_nlist"], r["code_size"] = cur.i64(), cur.i64()
    lists, caps = [], []
    for _ in range(r["l_nlist"]):
        size, cap = cur.i64(), cur.i64()
        ids = np.frombuffer(cur.take(8 * size), "<i8")
        codes = np.frombuffer(cur.take(size * r["code_size"]), np.uint8).reshape(size, r["code_size"])
        lists.append((ids, codes))
        caps.append(cap)
    r["lists"], r["list_caps"] = lists, caps
    assert cur.i64() == MAGIC_PQ
    r["p_d"], r["p_M"], r["ksub"] = cur.i64(), cur.i64(), cur.i64()
    r["p_trained"] = cur.byte()
    dsub = r["p_d"] // r["p_M"]
    r["pq"] = np.frombuffer(cur.take(4 * r["p_M"] * r["ksub"] * dsub), "<f4").reshape(
        r["p_M"], r["ksub"], dsub)
    return r


def parse_collection(buf: bytes) -> dict:
    """Checks the trailer and parses a collection snapshot."""
    payload, trailer = buf[:-16], buf[-16:]
    magic, value = struct.unpack("<QQ", trailer)
    assert magic == SUM_MAGIC and value == checksum(payload), "checksum"
    cur = _Cursor(payload)
    assert cur.i64() == COLL_MAGIC and cur.i64() == COLL_VERSION
    r = {"dimension": cur.i64()}
    r["name"] = cur.take(cur.i64()).decode()
    r["metric"] = cur.take(cur.i64()).decode()
    n = cur.i64()
    r["user_ids"] = np.frombuffer(cur.take(8 * n), "<i8")
    r["index"] = parse_index(cur)
    assert cur.pos == len(payload), "unexpected payload data"
    return r


def parse_index_file(buf: bytes) -> dict:
    cur = _Cursor(buf)
    r = parse_index(cur)
    assert cur.pos == len(buf)
    return r
