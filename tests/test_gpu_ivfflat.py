"""IVF-Flat on the GPU (through the C ABI) against the CPU oracle.

Bar: bit-exact ids and FP32 distances (the plain kernels keep the reference's
operation order: l2_distance_simd[8] and inner_product_simd[4],
index_ivf_flat.mojo:222-225); the stated fallback bound is 1e-5 relative.
Centroids come from the GPU k-means and are installed in the oracle, so the
comparison isolates list building and search; k-means parity has its own test.
"""
import os

import numpy as np
import pytest

from oracle import oracle as o

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
RTOL = 1e-5


def rint(rng, shape, lo=0, hi=256):
    return rng.integers(lo, hi, size=shape).astype(np.float32)


def assert_same(Dg, Ig, Do, Io):
    assert np.array_equal(Ig, Io), f"id mismatch at {np.argwhere(Ig != Io)[:5].tolist()}"
    assert np.array_equal(Dg, Do) or np.allclose(Dg, Do, rtol=RTOL, atol=0)


def pair(gpu, x, d, nlist, metric, ids=None, niter=5):
    """GPU index trained on x, oracle index with the same centroids, both
    holding the same rows."""
    g = gpu.IndexIVFFlat(d, nlist, metric)
    g.train(x, niter=niter)
    r = o.IVFFlatIndex(d, nlist, metric)
    r.set_centroids(g.centroids())
    if ids is None:
        g.add(x)
        r.add(x)
    else:
        g.add_with_ids(x, ids)
        r.add(x, ids=ids)
    return g, r


@pytest.mark.parametrize("d", [128, 96, 64, 32, 17, 8, 3])
@pytest.mark.parametrize("metric", [0, 1])
def test_ivfflat_matches_oracle(gpu, d, metric):
    rng = np.random.default_rng(100 + d + metric)
    x = rng.standard_normal((6000, d)).astype(np.float32)
    q = rng.standard_normal((41, d)).astype(np.float32)
    g, r = pair(gpu, x, d, 24, metric)
    for l in range(24):  # same lists, same append order
        gi, gr = g.get_list(l)
        ri, rr = r.list(l)
        assert np.array_equal(gi, ri) and np.array_equal(gr, rr)
    for nprobe in (1, 5, 24):
        g.nprobe = nprobe
        for k in (1, 10, 100):
            Dg, Ig = g.search(q, k)
            Do, Io = r.search(q, k, nprobe)
            assert_same(Dg, Ig, Do, Io)


@pytest.mark.parametrize("metric", [0, 1])
def test_ivfflat_ties_break_by_lower_id_with_unordered_ids(gpu, metric):
    # Few distinct values -> many exact distance ties; ids are a permutation,
    # so a list's append order is not its id order.
    rng = np.random.default_rng(7 + metric)
    x = rint(rng, (5000, 16), 0, 4)
    q = rint(rng, (30, 16), 0, 4)
    ids = rng.permutation(1 << 20)[:5000].astype(np.int64)
    g, r = pair(gpu, x, 16, 8, metric, ids=ids)
    for nprobe in (2, 8):
        g.nprobe = nprobe
        for k in (5, 50, 300):
            assert_same(*g.search(q, k), *r.search(q, k, nprobe))


@pytest.mark.parametrize("nlist", [1, 2, 3])
def test_ivfflat_lists_longer_than_one_chunk(gpu, nlist):
    # Lists of thousands of rows span several 1024-row chunks.
    rng = np.random.default_rng(30 + nlist)
    x = rint(rng, (9000, 32), 0, 64)
    q = rint(rng, (70, 32), 0, 64)
    g, r = pair(gpu, x, 32, nlist, 0, niter=3)
    assert g.list_sizes().max() > 2048
    g.nprobe = nlist
    for k in (1, 10, 1000):
        assert_same(*g.search(q, k), *r.search(q, k, nlist))


def test_ivfflat_exclusion_padding_and_repeated_adds(gpu):
    rng = np.random.default_rng(9)
    x = rng.standard_normal((3000, 64)).astype(np.float32)
    q = rng.standard_normal((25, 64)).astype(np.float32)
    g = gpu.IndexIVFFlat(64, 30)
    g.train(x, niter=4)
    r = o.IVFFlatIndex(64, 30)
    r.set_centroids(g.centroids())
    for part in np.array_split(x, 3):  # three adds: ids continue across calls
        g.add(part)
        r.add(part)
        g.nprobe = 3
        assert_same(*g.search(q, 10), *r.search(q, 10, 3))
    ex = (rng.random(3000) < 0.3).astype(np.uint8)
    assert_same(*g.search(q, 20, exclude=ex), *r.search(q, 20, 3, exclude=ex))
    # k beyond the rows of the probed lists: (1e38, -1) tail
    g.nprobe = 1
    Dg, Ig = g.search(q, 600)
    Do, Io = r.search(q, 600, 1)
    assert_same(Dg, Ig, Do, Io)
    assert (Ig == -1).any() and (Dg[Ig == -1] == np.float32(1e38)).all()
    # nprobe above nlist clamps to nlist
    g.nprobe = 1000
    assert_same(*g.search(q, 10), *r.search(q, 10, 30))


def test_ivfflat_reference_known_answers(gpu):
    # tests/test_ivf_flat.mojo:69-88: one vector, nlist 1 -> id 0, distance 0
    data = np.arange(4, dtype=np.float32)[None]
    g = gpu.IndexIVFFlat(4, 1)
    g.train(data)
    g.add(data)
    D, I = g.search(data, 1)
    assert I[0, 0] == 0 and D[0, 0] == 0.0
    # :91-137: k = 5 over one row with id 42 -> tail (1e38, -1)
    zero = np.zeros((1, 4), np.float32)
    g = gpu.IndexIVFFlat(4, 1)
    g.train(zero)
    g.add_with_ids(zero, np.array([42]))
    D, I = g.search(zero, 5)
    assert I[0].tolist() == [42, -1, -1, -1, -1]
    assert D[0, 0] == 0.0 and (D[0, 1:] == np.float32(1e38)).all()
    # :9-66: ids i * 10, nlist 10, nprobe 3
    rng = np.random.default_rng(101)
    x = (rng.random((1000, 16)) * 2 - 1).astype(np.float32)
    g = gpu.IndexIVFFlat(16, 10)
    g.train(x)
    g.add_with_ids(x, np.arange(1000) * 10)
    assert g.ntotal == 1000 and g.is_trained
    g.nprobe = 3
    D, I = g.search(x[:10], 5)
    assert (I[:, 0] == np.arange(10) * 10).all() and (np.diff(D, axis=1) >= 0).all()


def test_ivfflat_untrained_and_empty_pad(gpu):
    g = gpu.IndexIVFFlat(8, 4)
    rows = np.load(os.path.join(GOLDEN, "ivfpq_rows.npz"))["rows"]
    g.add(rows)  # dropped while untrained (index_ivf_flat.mojo:102-104)
    assert g.ntotal == 0
    D, I = g.search(rows[:2], 3)
    assert (I == -1).all() and (D == np.float32(1e38)).all()
    g.train(rows)
    D, I = g.search(rows[:2], 3)
    assert (I == -1).all() and (D == np.float32(1e38)).all()
    with pytest.raises(gpu.MojoVecError, match="ids"):
        g.add_with_ids(rows[:2], np.array([-1, 3]))


def test_ivfflat_kmeans_matches_oracle_on_integer_data(gpu):
    # On integer data every k-means sum is exact in FP32 and FP64, so the GPU
    # coarse centroids equal the oracle's KMeans(d, nlist, niter) restatement.
    rng = np.random.default_rng(12)
    x = rint(rng, (4000, 32), 0, 32)
    g = gpu.IndexIVFFlat(32, 16)
    g.train(x, niter=6)
    r = o.IVFFlatIndex(32, 16)
    r.train(x, niter=6)
    assert np.allclose(g.centroids(), r.centroids(), rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("metric", [0, 1])
def test_ivfflat_full_probe_equals_exact_flat_at_scale(gpu, metric):
    # nprobe = nlist visits every row: on integer data in [0, 255] every FP32
    # summation order is exact, so the result equals the exact Flat search
    # (GPU IndexFlat) over the same rows, bit for bit.
    rng = np.random.default_rng(40 + metric)
    x = rint(rng, (200000, 128))
    q = rint(rng, (500, 128))
    g = gpu.IndexIVFFlat(128, 256, metric)
    g.train(x[:50000], niter=3)
    g.add(x)
    g.nprobe = 256
    f = gpu.IndexFlat(128, metric)
    f.add(x)
    Dg, Ig = g.search(q, 10)
    Df, If = f.search(q, 10)
    assert np.array_equal(Ig, If) and np.array_equal(Dg, Df)
