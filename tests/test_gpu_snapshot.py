"""IVF-PQ snapshots on the GPU: the reference's save/load checks
(python/tests/test_binding_ivfpq.py:123-167, tests/test_ivf_pq.mojo:155-206),
files written by the oracle's independent restatement of the format loaded
into HBM and searched, and files written by the library parsed back by it."""
import numpy as np
import pytest

from oracle import oracle as o
from oracle import snapshot as osnap

pytestmark = pytest.mark.gpu

DIMENSION = 8
COUNT = 256


def vectors():
    return [
        [((row * 17 + column * 29) % 251) / 251.0 + (2.0 if column == row % DIMENSION else 0.0)
         for column in range(DIMENSION)]
        for row in range(COUNT)
    ]


def clustered(seed, n, d, nclusters=30):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0, 128, (nclusters, d))
    x = centres[rng.integers(0, nclusters, n)] + rng.normal(0, 12.0, (n, d))
    return np.rint(np.clip(x, 0, 255)).astype(np.float32)


def test_collection_save_load_round_trip(gpu, tmp_path):
    path = tmp_path / "index.ivfpq"
    rows = vectors()
    c = gpu.IVFPQCollection(DIMENSION, nlist=8, pq_subvectors=2, nprobe=4, metric="cosine",
                            name="persistent")
    c.add(list(range(COUNT)), rows)
    expected = c.query(rows[0], n_results=3)
    c.save(path)
    loaded = gpu.IVFPQCollection.load(path)
    assert loaded.name() == "persistent" and loaded.metric() == "cosine"
    assert loaded.count() == COUNT and loaded.nprobe() == 4 and loaded.is_trained()
    assert loaded.query(rows[0], n_results=3) == expected
    payload = bytearray(path.read_bytes())
    payload[len(payload) // 2] ^= 0x01
    path.write_bytes(bytes(payload))
    with pytest.raises(Exception, match="checksum"):
        gpu.IVFPQCollection.load(path)


def test_untrained_save_load(gpu, tmp_path):
    path = tmp_path / "empty.ivfpq"
    c = gpu.IVFPQCollection(DIMENSION, nlist=8, pq_subvectors=2, nprobe=4, metric="ip",
                            name="empty")
    c.save(path)
    loaded = gpu.IVFPQCollection.load(path)
    assert not loaded.is_trained() and loaded.count() == 0
    assert loaded.metric() == "ip" and loaded.nprobe() == 4 and loaded.name() == "empty"
    # the loaded collection trains on its first add, as a fresh one does
    rows = vectors()
    loaded.add(list(range(COUNT)), rows)
    assert loaded.is_trained() and loaded.count() == COUNT


def _oracle_index(d=32, nlist=16, M=8, n=3000):
    x = clustered(11, n, d)
    oi = o.IVFPQIndex(d, nlist, M, 0)
    oi.train(x, niter_coarse=4, niter_pq=4)
    oi.add(x)
    coarse, pq = oi.codebooks()
    lists = [oi.list(l) for l in range(nlist)]
    return x, oi, coarse, pq, lists


def test_oracle_written_file_loads_and_searches(gpu, tmp_path):
    d, nlist, M, n, nprobe = 32, 16, 8, 3000, 6
    x, oi, coarse, pq, lists = _oracle_index(d, nlist, M, n)
    user_ids = 7000 + 3 * np.arange(n)
    index = osnap.index_bytes(d, nlist, M, nprobe, n, True, 0, coarse, lists, pq)
    path = tmp_path / "oracle.ivfpq"
    data = osnap.collection_bytes(d, "interop", "l2", user_ids, index)
    path.write_bytes(data)
    c = gpu.IVFPQCollection.load(path)
    assert c.count() == n and c.nprobe() == nprobe and c.name() == "interop"
    q = clustered(12, 25, d)
    r = c.query_numpy(q, n_results=10)
    Do, Io = oi.search(q, 10, nprobe)
    assert np.array_equal(np.asarray(r["ids"]), user_ids[Io])
    assert np.array_equal(np.asarray(r["distances"], np.float32), Do)
    # load -> save reproduces the file byte for byte
    again = tmp_path / "again.ivfpq"
    c.save(again)
    assert again.read_bytes() == data


def test_loaded_capacities_are_kept_and_grow(gpu, tmp_path):
    d, nlist, M, n = 32, 16, 8, 3000
    x, oi, coarse, pq, lists = _oracle_index(d, nlist, M, n)
    caps = [0 if ids.size == 0 else max(16, ids.size + 5 + l) for l, (ids, _) in enumerate(lists)]
    index = osnap.index_bytes(d, nlist, M, 3, n, True, 0, coarse, lists, pq,
                              coarse_capacity=2000, list_caps=caps)
    data = osnap.collection_bytes(d, "caps", "l2", np.arange(n), index)
    path = tmp_path / "caps.ivfpq"
    path.write_bytes(data)
    c = gpu.IVFPQCollection.load(path)
    c.save(tmp_path / "same.ivfpq")
    assert (tmp_path / "same.ivfpq").read_bytes() == data
    extra = clustered(13, 500, d)
    c.add_numpy(np.arange(n, n + 500, dtype=np.int64), extra)
    c.save(tmp_path / "grown.ivfpq")
    r = osnap.parse_collection((tmp_path / "grown.ivfpq").read_bytes())
    ix = r["index"]
    assert ix["q_capacity"] == 2000 and ix["ntotal"] == n + 500
    for l, (ids, codes) in enumerate(ix["lists"]):
        assert ix["list_caps"][l] == osnap.list_capacity(ids.size, caps[l])
        assert np.array_equal(ids[: lists[l][0].size], lists[l][0])  # append order kept
    assert np.array_equal(r["user_ids"], np.arange(n + 500))


def test_library_written_file_parses(gpu, tmp_path):
    rows = np.asarray(vectors(), np.float32)
    c = gpu.IVFPQCollection(DIMENSION, nlist=8, pq_subvectors=2, nprobe=3, metric="l2", name="w")
    c.add_numpy(np.arange(500, 500 + COUNT, dtype=np.int64), rows)
    c.save(tmp_path / "w.ivfpq")
    buf = (tmp_path / "w.ivfpq").read_bytes()
    r = osnap.parse_collection(buf)
    assert r["name"] == "w" and r["metric"] == "l2" and r["dimension"] == DIMENSION
    assert np.array_equal(r["user_ids"], np.arange(500, 500 + COUNT))
    ix = r["index"]
    assert (ix["d"], ix["nlist"], ix["M"], ix["nprobe"], ix["ntotal"]) == (DIMENSION, 8, 2, 3, COUNT)
    assert ix["trained"] == 1 and ix["metric"] == 0 and ix["q_capacity"] == 1024
    assert not ix["coarse_tail"].any()
    coarse, pq = c.index().codebooks()
    assert np.array_equal(ix["coarse"], coarse) and np.array_equal(ix["pq"], pq)
    gi = c.index()
    for l in range(8):
        ids, codes = gi.get_list(l)
        assert np.array_equal(ix["lists"][l][0], ids) and np.array_equal(ix["lists"][l][1], codes)
        assert ix["list_caps"][l] == osnap.list_capacity(ids.size)
    # the oracle's writer produces the same bytes from the same content
    index = osnap.index_bytes(DIMENSION, 8, 2, 3, COUNT, True, 0, coarse,
                              [gi.get_list(l) for l in range(8)], pq)
    assert osnap.collection_bytes(DIMENSION, "w", "l2", np.arange(500, 500 + COUNT), index) == buf


def test_index_level_save_load(gpu, tmp_path):
    d, nlist, M = 32, 16, 8
    x = clustered(14, 2000, d)
    gi = gpu.IndexIVFPQ(d, nlist, M)
    gi.train(x, niter_coarse=4, niter_pq=4)
    gi.add(x)
    gi.nprobe = 5
    q = clustered(15, 20, d)
    D0, I0 = gi.search(q, 10)
    gi.save(tmp_path / "ix.bin")
    ld = gpu.IndexIVFPQ.load(tmp_path / "ix.bin")
    assert (ld.d, ld.nlist, ld.M, ld.metric_type, ld.nprobe, ld.ntotal) == (d, nlist, M, 0, 5, 2000)
    D1, I1 = ld.search(q, 10)
    assert np.array_equal(I0, I1) and np.array_equal(D0, D1)
    r = osnap.parse_index_file((tmp_path / "ix.bin").read_bytes())
    assert r["ntotal"] == 2000 and r["nprobe"] == 5
    with pytest.raises(Exception, match="k_factor"):
        ld.search(q, 10, k_factor=2)  # snapshots carry no FP32 rows


# ---------------------------------------------------------------- IVF-Flat
@pytest.mark.parametrize("metric,qmetric", [(0, 0), (1, 1), (1, 0)])
def test_ivfflat_save_load_round_trip(gpu, tmp_path, metric, qmetric):
    d, nlist = 32, 16
    x = clustered(21, 3000, d)
    gi = gpu.IndexIVFFlat(d, nlist, metric, quantizer_metric=qmetric)
    gi.train(x, niter=4)
    gi.add(x)
    gi.nprobe = 5
    q = clustered(22, 25, d)
    D0, I0 = gi.search(q, 10)
    gi.save(tmp_path / "ivf.bin")
    ld = gpu.IndexIVFFlat.load(tmp_path / "ivf.bin")
    assert (ld.d, ld.nlist, ld.metric_type, ld.quantizer_metric) == (d, nlist, metric, qmetric)
    assert (ld.nprobe, ld.ntotal, ld.is_trained) == (5, 3000, True)
    D1, I1 = ld.search(q, 10)
    assert np.array_equal(I0, I1) and np.array_equal(D0, D1)
    # the oracle parses the file and, given its content, searches the same
    r = osnap.parse_ivfflat((tmp_path / "ivf.bin").read_bytes())
    assert r["metric"] == metric and r["q_metric"] == qmetric and r["code_size"] == 4 * d
    assert r["q_capacity"] == 1024 and not r["coarse_tail"].any()
    oi = o.IVFFlatIndex(d, nlist, metric, quantizer_metric=qmetric)
    oi.set_centroids(r["coarse"])
    for ids, vecs in r["lists"]:
        oi.add(np.asarray(vecs), ids=np.asarray(ids))  # re-adding keeps each row's list
    Do, Io = oi.search(q, 10, 5)
    assert np.array_equal(I0, Io) and np.array_equal(D0, Do)
    # load -> save writes the same bytes
    ld.save(tmp_path / "again.bin")
    assert (tmp_path / "again.bin").read_bytes() == (tmp_path / "ivf.bin").read_bytes()


def test_ivfflat_oracle_written_file_loads_and_searches(gpu, tmp_path):
    d, nlist = 16, 8
    rng = np.random.default_rng(23)
    x = rng.standard_normal((900, d)).astype(np.float32)
    oi = o.IVFFlatIndex(d, nlist)
    oi.train(x, niter=3)
    ids = rng.permutation(5000)[:900]
    oi.add(x, ids=ids)
    lists = [oi.list(l) for l in range(nlist)]
    caps = [osnap.list_capacity(len(li[0])) * 2 for li in lists]  # a grown capacity
    data = osnap.ivfflat_bytes(d, nlist, 3, 900, True, 0, oi.centroids(), lists,
                               coarse_capacity=4096, list_caps=caps)
    (tmp_path / "o.bin").write_bytes(data)
    gi = gpu.IndexIVFFlat.load(tmp_path / "o.bin")
    q = rng.standard_normal((30, d)).astype(np.float32)
    Dg, Ig = gi.search(q, 20)
    Do, Io = oi.search(q, 20, 3)
    assert np.array_equal(Ig, Io) and np.array_equal(Dg, Do)
    gi.save(tmp_path / "g.bin")
    assert (tmp_path / "g.bin").read_bytes() == data  # capacities kept
    gi.add(x[:40])  # sequential ids continue from ntotal; capacities grow by the rules
    gi.save(tmp_path / "g2.bin")
    r = osnap.parse_ivfflat((tmp_path / "g2.bin").read_bytes())
    assert r["ntotal"] == 940
    for l in range(nlist):
        assert r["list_caps"][l] == osnap.list_capacity(len(r["lists"][l][0]), caps[l])


def test_ivfflat_untrained_save_load(gpu, tmp_path):
    gi = gpu.IndexIVFFlat(8, 4)
    gi.save(tmp_path / "u.bin")
    r = osnap.parse_ivfflat((tmp_path / "u.bin").read_bytes())
    assert r["trained"] == 0 and r["q_rows"] == 0 and r["q_capacity"] == 1024
    ld = gpu.IndexIVFFlat.load(tmp_path / "u.bin")
    assert not ld.is_trained and ld.ntotal == 0
    D, I = ld.search(np.zeros((2, 8), np.float32), 3)
    assert (I == -1).all() and (D == np.float32(1e38)).all()
