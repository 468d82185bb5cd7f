"""IVFPQCollection drop-in behaviour on the GPU: the checks of MojoVec's
python/tests/test_binding_ivfpq.py:26-198 and tests/test_ivf_pq.mojo:48-152,
run against mojovec_amd."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMENSION = 8
COUNT = 256


def vectors():
    return [
        [((row * 17 + column * 29) % 251) / 251.0 + (2.0 if column == row % DIMENSION else 0.0)
         for column in range(DIMENSION)]
        for row in range(COUNT)
    ]


def test_lifecycle_and_stats(gpu):
    c = gpu.IVFPQCollection(DIMENSION, nlist=8, pq_subvectors=2, nprobe=4, metric="l2",
                            name="compressed")
    rows = vectors()
    ids = list(range(10_000, 10_000 + COUNT))
    assert not c.is_trained()
    c.train(rows)
    c.add(ids, rows)
    assert c.is_trained() and c.count() == COUNT
    assert c.name() == "compressed" and c.metric() == "l2" and c.dimension() == DIMENSION
    assert c.nlist() == 8 and c.pq_subvectors() == 2 and c.nprobe() == 4
    assert c.stats() == {"count": COUNT, "dimension": DIMENSION, "nlist": 8, "M": 2, "nprobe": 4,
                         "trained": True, "metric": "l2", "code_size_bytes": 2}
    r = c.query([rows[0]], n_results=3)
    assert r["ids"][0][0] == 10_000 and r["distances"][0][0] < 1e-4
    assert r["metadatas"] == [] and r["documents"] == [] and r["scores"] == []
    c.set_nprobe(8)
    assert c.nprobe() == 8


def test_auto_training_validation_and_duplicates(gpu):
    c = gpu.IVFPQCollection(DIMENSION, nlist=8, pq_subvectors=2)
    rows = vectors()
    c.add(list(range(COUNT)), rows)
    assert c.is_trained()
    with pytest.raises(Exception, match="already exists"):
        c.add([0], [rows[0]])
    with pytest.raises(Exception, match="Duplicate IDs"):
        c.add([1000, 1000], [rows[0], rows[1]])
    with pytest.raises(Exception, match="nprobe"):
        c.set_nprobe(9)
    with pytest.raises(ValueError, match="expected 8"):
        c.add([999], [[1.0, 2.0]])
    with pytest.raises(Exception, match="at least max"):
        gpu.IVFPQCollection(DIMENSION, nlist=8, pq_subvectors=2).train(rows[:255])
    with pytest.raises(Exception, match="finite"):
        gpu.IVFPQCollection(DIMENSION, nlist=8, pq_subvectors=2).train(
            [[math.nan] * DIMENSION for _ in range(COUNT)])
    with pytest.raises(Exception, match="n_results"):
        c.query(rows[0], n_results=0)


@pytest.mark.parametrize("metric", ["cosine", "ip"])
def test_similarity_metrics(gpu, metric):
    rows = vectors()
    query = list(rows[0])
    if metric == "ip":
        rows[0] = [v * 20.0 for v in rows[0]]
        query = rows[0]
    c = gpu.IVFPQCollection(DIMENSION, nlist=8, pq_subvectors=2, nprobe=8, metric=metric)
    c.add(list(range(COUNT)), rows)
    r = c.query(query, n_results=1)
    assert r["ids"] == [[0]]
    if metric == "cosine":
        assert r["distances"][0][0] < 1e-4
    else:
        assert r["distances"][0][0] < 0.0


def test_cosine_rejects_zero_vectors(gpu):
    c = gpu.IVFPQCollection(DIMENSION, nlist=8, pq_subvectors=2, metric="cosine")
    rows = vectors()
    rows[3] = [0.0] * DIMENSION
    with pytest.raises(Exception, match="non-zero"):
        c.add(list(range(COUNT)), rows)


def test_numpy_fast_path(gpu):
    rows = np.asarray(vectors(), dtype=np.float32)
    ids = np.arange(COUNT, dtype=np.int64)
    c = gpu.IVFPQCollection(DIMENSION, nlist=8, pq_subvectors=2, nprobe=8)
    c.train_numpy(rows)
    c.add_numpy(ids, rows)
    r = c.query_numpy(rows[:2], n_results=2)
    assert r["ids"][0][0] == 0 and r["ids"][1][0] == 1
    with pytest.raises(TypeError, match="float32"):
        c.query_numpy(rows.astype(np.float64))
    with pytest.raises(TypeError, match="int64"):
        gpu.IVFPQCollection(DIMENSION, nlist=8, pq_subvectors=2).add_numpy(ids.astype(np.int32), rows)


def test_k_factor_rerank_recovers_exact_neighbours(gpu):
    rng = np.random.default_rng(3)
    x = rng.standard_normal((5000, 32)).astype(np.float32)
    q = rng.standard_normal((50, 32)).astype(np.float32)
    c = gpu.IVFPQCollection(32, nlist=16, pq_subvectors=8, nprobe=16)
    c.add_numpy(np.arange(5000, dtype=np.int64), x)
    d2 = ((q[:, None, :].astype(np.float64) - x[None]) ** 2).sum(-1)
    gt = np.argsort(d2, 1)[:, :10]
    rec = lambda r: np.mean([len(set(a) & set(b)) / 10 for a, b in zip(gt.tolist(), r["ids"])])
    plain = rec(c.query_numpy(q, 10))
    reranked = rec(c.query_numpy(q, 10, k_factor=8))
    assert reranked > plain and reranked >= 0.9


def test_repr(gpu):
    c = gpu.IVFPQCollection(8, nlist=4, pq_subvectors=2, name="docs")
    assert "IVFPQCollection" in repr(c) and "name='docs'" in repr(c)
