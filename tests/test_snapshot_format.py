"""IVF-PQ snapshot format on the host: the library's checksum against the
oracle's restatement (mojovec/io/snapshot_file.mojo:19-59), and the load
validation of CollectionIVFPQ.load (mojovec/api/collection_ivfpq.mojo:372-476,
mojovec/io/serialization.mojo) on crafted files. Every rejected file fails
before the library touches a device, so these run without a GPU."""
import ctypes
import struct

import numpy as np
import pytest

from oracle import snapshot as osnap

from mojovec_amd import _lib


def lib_checksum(payload: bytes) -> int:
    buf = ctypes.create_string_buffer(payload, len(payload))
    out = ctypes.c_uint64()
    _lib.check(_lib.load().mvg_snapshot_checksum(buf, len(payload), ctypes.byref(out)))
    return out.value


@pytest.mark.parametrize("n", [0, 1, 7, 8, 9, 15, 16, 17, 100, 4096 + 3])
def test_checksum_matches_oracle_small(n):
    payload = np.random.default_rng(n).integers(0, 256, n, dtype=np.uint8).tobytes()
    assert lib_checksum(payload) == osnap.checksum(payload)


def test_checksum_chunk_boundary():
    # 8 MiB chunks: the second chunk restarts the length term.
    n = osnap.SUM_CHUNK + 13
    payload = np.random.default_rng(5).integers(0, 256, n, dtype=np.uint8).tobytes()
    got = lib_checksum(payload)
    assert got == osnap.checksum(payload)
    one_chunk = osnap.checksum_chunk(osnap.SUM_SEED, payload)  # wrong chunking
    assert got != one_chunk


def test_checksum_properties():
    # Empty payload: the seed itself (the streaming loop never runs).
    assert lib_checksum(b"") == osnap.SUM_SEED
    # Any single-bit flip of a small payload changes the sum.
    base = bytes(range(40))
    s0 = lib_checksum(base)
    for i in range(len(base)):
        for bit in (0, 7):
            flipped = bytearray(base)
            flipped[i] ^= 1 << bit
            assert lib_checksum(bytes(flipped)) != s0
    # Trailing zero bytes are not ignored (length enters every chunk).
    assert lib_checksum(base + b"\0") != s0


# ---------------------------------------------------------------- load checks
D, NLIST, M = 8, 4, 2


def make_index(**over):
    rng = np.random.default_rng(0)
    ids = np.arange(10)
    lists = [(ids[l::NLIST], rng.integers(0, 256, (len(ids[l::NLIST]), M), dtype=np.uint8))
             for l in range(NLIST)]
    args = dict(d=D, nlist=NLIST, M=M, nprobe=2, ntotal=10, trained=True, metric=0,
                coarse=rng.standard_normal((NLIST, D)), lists=lists,
                pq=rng.standard_normal((M, 256, D // M)))
    args.update(over)
    return osnap.index_bytes(**args)


def make_file(tmp_path, index=None, name="c", metric="l2", user_ids=None, dimension=D):
    index = make_index() if index is None else index
    user_ids = np.arange(100, 110) if user_ids is None else user_ids
    data = osnap.collection_bytes(dimension, name, metric, user_ids, index)
    p = tmp_path / "snap.ivfpq"
    p.write_bytes(data)
    return p, data


def reseal(tmp_path, payload: bytes):
    p = tmp_path / "snap.ivfpq"
    p.write_bytes(payload + struct.pack("<QQ", osnap.SUM_MAGIC, osnap.checksum(payload)))
    return p


def load_error(path) -> str:
    h = ctypes.c_void_p()
    rc = _lib.load().mvg_coll_load(str(path).encode(), ctypes.byref(h))
    assert rc != _lib.MVG_OK, "crafted snapshot was accepted"
    return _lib.load().mvg_last_error().decode()


def header_len(name="c", metric="l2", n_ids=10):
    return 8 * 3 + 8 + len(name) + 8 + len(metric) + 8 + 8 * n_ids


def patch_i64(buf: bytes, pos: int, value: int) -> bytes:
    b = bytearray(buf)
    b[pos:pos + 8] = struct.pack("<q", value)
    return bytes(b)


def test_trailer_checks(tmp_path):
    p = tmp_path / "short.ivfpq"
    p.write_bytes(b"\0" * 15)
    assert load_error(p) == "IVF-PQ snapshot checksum trailer is missing."
    p, data = make_file(tmp_path)
    bad = bytearray(data)
    bad[-16] ^= 1
    p.write_bytes(bytes(bad))
    assert load_error(p) == "Invalid snapshot checksum trailer."
    bad = bytearray(data)
    bad[len(data) // 2] ^= 0x01  # python/tests/test_binding_ivfpq.py:144-148
    p.write_bytes(bytes(bad))
    assert "checksum" in load_error(p)
    assert load_error(p) == "Snapshot checksum mismatch."


def test_header_checks(tmp_path):
    _, data = make_file(tmp_path)
    payload = data[:-16]
    assert load_error(reseal(tmp_path, patch_i64(payload, 0, 7))) == \
        "Invalid CollectionIVFPQ snapshot magic."
    assert load_error(reseal(tmp_path, patch_i64(payload, 8, 2))) == \
        "Unsupported CollectionIVFPQ snapshot version."
    assert load_error(reseal(tmp_path, patch_i64(payload, 16, 0))) == \
        "dimension must be between 1 and 65536."
    p, _ = make_file(tmp_path, metric="dot")
    assert load_error(p) == 'metric must be "l2", "cosine", or "ip".'
    assert load_error(reseal(tmp_path, payload + b"\0")) == \
        "IVF-PQ snapshot contains unexpected payload data."
    # The stream is read from the file, trailer included (as the reference
    # does): a payload short by more than the trailer runs out of data.
    assert load_error(reseal(tmp_path, payload[:-40])) == "Unexpected end of serialized data."


def test_index_stream_checks(tmp_path):
    _, data = make_file(tmp_path)
    payload = data[:-16]
    h = header_len()
    assert load_error(reseal(tmp_path, patch_i64(payload, h, 1))) == "Invalid magic for IndexIVFPQ"
    # nprobe > nlist: check_size_limit (serialization.mojo:40-43)
    assert load_error(reseal(tmp_path, patch_i64(payload, h + 32, NLIST + 1))).startswith(
        "Security Error")
    assert load_error(reseal(tmp_path, patch_i64(payload, h + 32, 0))) == "Invalid IndexIVFPQ nprobe."
    assert load_error(reseal(tmp_path, patch_i64(payload, h + 24, 3))) == \
        "Invalid IndexIVFPQ subvector count."
    b = bytearray(payload)
    b[h + 48] = 2  # is_trained Bool
    assert load_error(reseal(tmp_path, bytes(b))) == "Invalid serialized Bool value."
    assert load_error(reseal(tmp_path, patch_i64(payload, h + 49, 2))) == "Invalid IndexIVFPQ metric."
    # coarse quantizer header: ntotal must equal nlist when trained
    q = h + 57
    assert load_error(reseal(tmp_path, patch_i64(payload, q, 5))) == "Invalid magic for IndexFlat"
    assert load_error(reseal(tmp_path, patch_i64(payload, q + 16, NLIST - 1))) == \
        "IndexIVFPQ coarse quantizer header mismatch."


def test_list_and_pq_checks(tmp_path):
    # the PQ training flag must match the index header
    _, data = make_file(tmp_path)
    payload = bytearray(data[:-16])
    pq_flag = len(payload) - 4 * D * 256 - 1
    payload[pq_flag] = 0
    assert load_error(reseal(tmp_path, bytes(payload))) == "IndexIVFPQ training state mismatch."
    ksub_pos = pq_flag - 8
    assert load_error(reseal(tmp_path, patch_i64(bytes(payload), ksub_pos, 128))) == \
        "ProductQuantizer shape does not match its index header."
    h = header_len()
    inv = h + 57 + 40 + 4 * 1024 * D  # after the IndexFlat (capacity 1024)
    assert load_error(reseal(tmp_path, patch_i64(data[:-16], inv + 16, M + 1))) == \
        "Inverted-list shape does not match its index header."
    assert load_error(reseal(tmp_path, patch_i64(data[:-16], inv, 3))) == \
        "Invalid magic for ArrayInvertedLists"


def test_id_checks(tmp_path):
    p, _ = make_file(tmp_path, user_ids=np.arange(100, 109))  # 9 ids, ntotal 10
    assert load_error(p) == "IVF-PQ snapshot dimensions or record count mismatch."
    p, _ = make_file(tmp_path, metric="ip")  # index metric is L2
    assert load_error(p) == "IVF-PQ snapshot metric mismatch."
    rng = np.random.default_rng(1)

    def with_ids(list0_ids):
        lists = [(np.asarray(list0_ids), rng.integers(0, 256, (len(list0_ids), M), dtype=np.uint8))]
        lists += [(np.arange(0), np.zeros((0, M), np.uint8))] * (NLIST - 1)
        return make_index(lists=lists)

    p, _ = make_file(tmp_path, index=with_ids([0, 1, 2, 3, 4, 5, 6, 7, 8, 10]))
    assert load_error(p) == "IVF-PQ snapshot contains an invalid internal ID."
    p, _ = make_file(tmp_path, index=with_ids([0, 1, 2, 3, 4, 5, 6, 7, 8, 8]))
    assert load_error(p) == "IVF-PQ snapshot contains a duplicate internal ID."
    p, _ = make_file(tmp_path, index=with_ids([0, 1, 2, 3, 4, 5, 6, 7, 8]))
    assert load_error(p) == "IVF-PQ inverted-list count does not match its IDs."
    p, _ = make_file(tmp_path, user_ids=np.array([5, 1, 2, 3, 4, 5, 6, 7, 8, 9]))
    assert load_error(p) == "IVF-PQ snapshot contains duplicate IDs."


def test_oracle_capacity_rules():
    # inverted_lists.mojo:119-130, appended one entry at a time from empty
    cap = 0
    for size in range(1, 200):
        if size > cap:
            cap = max(16, size) if cap == 0 else max(2 * cap, size)
        assert osnap.list_capacity(size) == cap
    assert osnap.list_capacity(0) == 0
    assert osnap.flat_capacity(1024) == 1024 and osnap.flat_capacity(1025) == 2048
    assert osnap.flat_capacity(16384) == 16384


# ---------------------------------------------------------------- IVF-Flat files
def make_ivfflat(**over):
    rng = np.random.default_rng(3)
    ids = np.arange(10)
    lists = [(ids[l::NLIST], rng.standard_normal((len(ids[l::NLIST]), D))) for l in range(NLIST)]
    args = dict(d=D, nlist=NLIST, nprobe=2, ntotal=10, trained=True, metric=0,
                coarse=rng.standard_normal((NLIST, D)), lists=lists)
    args.update(over)
    return osnap.ivfflat_bytes(**args)


def ivfflat_error(tmp_path, data: bytes) -> str:
    p = tmp_path / "ix.ivfflat"
    p.write_bytes(data)
    h = ctypes.c_void_p()
    rc = _lib.load().mvg_ivfflat_load(str(p).encode(), ctypes.byref(h))
    assert rc != _lib.MVG_OK, "crafted IVF-Flat file was accepted"
    return _lib.load().mvg_last_error().decode()


def test_ivfflat_load_checks(tmp_path):
    # read_index_ivf_flat (serialization.mojo:995-1020) and its parts
    data = make_ivfflat()
    assert ivfflat_error(tmp_path, patch_i64(data, 0, 1)) == "Invalid magic for IndexIVFFlat"
    assert ivfflat_error(tmp_path, patch_i64(data, 8, 0)) == "dimension must be between 1 and 65536."
    assert ivfflat_error(tmp_path, patch_i64(data, 16, 1000001)).startswith("Security Error")
    assert ivfflat_error(tmp_path, patch_i64(data, 24, 1000001)).startswith("Security Error")
    assert ivfflat_error(tmp_path, patch_i64(data, 32, -1)).startswith("Security Error")
    b = bytearray(data)
    b[40] = 3
    assert ivfflat_error(tmp_path, bytes(b)) == "Invalid serialized Bool value."
    q = 49  # the quantizer IndexFlat
    assert ivfflat_error(tmp_path, patch_i64(data, q, 5)) == "Invalid magic for IndexFlat"
    inv = q + 40 + 4 * 1024 * D
    assert ivfflat_error(tmp_path, patch_i64(data, inv, 9)) == "Invalid magic for ArrayInvertedLists"
    assert ivfflat_error(tmp_path, patch_i64(data, inv + 16, 4 * D + 1)) == \
        "Inverted-list shape does not match its index header."
    assert ivfflat_error(tmp_path, data[:-3]) == "Unexpected end of serialized data."
    # checks beyond the reference reader: a quantizer contradicting the
    # header, nprobe 0, ids the device lists cannot hold
    assert ivfflat_error(tmp_path, patch_i64(data, q + 16, NLIST - 1)) == \
        "IndexIVFFlat coarse quantizer header mismatch."
    assert ivfflat_error(tmp_path, patch_i64(data, 24, 0)) == "Invalid IndexIVFFlat nprobe."
    rng = np.random.default_rng(4)
    lists = [(np.array([1 << 31]), rng.standard_normal((1, D)))] + \
        [(np.arange(0), np.zeros((0, D)))] * (NLIST - 1)
    assert ivfflat_error(tmp_path, make_ivfflat(lists=lists)) == \
        "IVF-Flat snapshot contains an id outside [0, 2^31 - 1]."


def test_ivfflat_oracle_layout_sizes():
    # header 49 bytes, IndexFlat 40 + 4 * capacity * d, invlists 24 + per
    # list 16 + 8 * size + 4 * d * size
    data = make_ivfflat()
    assert len(data) == 49 + 40 + 4 * 1024 * D + 24 + NLIST * 16 + 10 * (8 + 4 * D)
    r = osnap.parse_ivfflat(data)
    assert r["code_size"] == 4 * D and r["q_capacity"] == 1024 and r["list_caps"] == [16] * NLIST
