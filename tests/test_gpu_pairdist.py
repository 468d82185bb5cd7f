"""Batched distance service (the Flat and SQ8 distance computers a CPU graph
search calls) on the GPU against the oracle: bit-exact values for both FP32
summation orders (distance and distance_batch4, index_flat.mojo:60-90), the
symmetric form (:92-99) and the SQ8 code distances (index_flat_sq8.mojo:80-177).
"""
import numpy as np
import pytest

from oracle import oracle as o

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("d", [3, 7, 8, 17, 32, 45, 96, 128, 200])
@pytest.mark.parametrize("metric", [0, 1])
def test_flat_distances_bit_exact(gpu, d, metric):
    rng = np.random.default_rng(1000 + d + metric)
    base = (rng.standard_normal((3000, d)) * rng.choice([0.1, 1.0, 30.0], (3000, 1))).astype(np.float32)
    q = rng.standard_normal((37, d)).astype(np.float32)
    ids = rng.integers(0, 3000, (37, 53))
    ix = gpu.IndexFlat(d, metric)
    ix.add(base)
    for order, name in ((0, "distance"), (1, "batch4")):
        got = ix.distances(q, ids, order=name)
        want = o.flat_distances(base, q, ids, metric, order)
        assert np.array_equal(got, want), (order, np.argwhere(got != want)[:5].tolist())
    a, b = rng.integers(0, 3000, 999), rng.integers(0, 3000, 999)
    assert np.array_equal(ix.pair_distances(a, b), o.flat_pair_distances(base, a, b, metric))


@pytest.mark.parametrize("d", [3, 24, 100, 128, 768])
@pytest.mark.parametrize("metric", [0, 1])
def test_sq8_distances_bit_exact(gpu, d, metric):
    rng = np.random.default_rng(2000 + d + metric)
    x = rng.standard_normal((2500, d)).astype(np.float32)
    q = rng.standard_normal((21, d)).astype(np.float32) * 1.3  # some query values clamp
    ids = rng.integers(0, 2500, (21, 40))
    gx = gpu.IndexFlatSQ8(d, metric)
    ox = o.SQ8Index(d, metric)
    for part in (x[:1000] * 0.5, x[1000:]):  # the second add widens the range (re-encode)
        gx.add(part)
        ox.add(part)
    assert np.array_equal(gx.distances(q, ids), ox.distances(q, ids))
    a, b = rng.integers(0, 2500, 700), rng.integers(0, 2500, 700)
    assert np.array_equal(gx.pair_distances(a, b), ox.pair_distances(a, b))


def test_distance_service_large_batch_and_checks(gpu):
    rng = np.random.default_rng(5)
    base = rng.standard_normal((20000, 64)).astype(np.float32)
    q = rng.standard_normal((500, 64)).astype(np.float32)
    ids = rng.integers(0, 20000, (500, 300))
    ix = gpu.IndexFlat(64)
    ix.add(base)
    got = ix.distances(q, ids, order="batch4")
    assert np.array_equal(got, o.flat_distances(base, q, ids, 0, 1))
    # empty batches are no-ops
    assert ix.distances(q[:0], ids[:0]).shape == (0, 300)
    assert ix.pair_distances([], []).shape == (0,)
    with pytest.raises(gpu.MojoVecError, match="ids"):
        ix.distances(q[:1], np.array([[0, 20000]]))
    with pytest.raises(gpu.MojoVecError, match="ids"):
        ix.pair_distances([-1], [0])
    with pytest.raises(KeyError):
        ix.distances(q[:1], np.array([[0]]), order="other")
    sq = gpu.IndexFlatSQ8(64, 1)
    sq.add(base)
    with pytest.raises(gpu.MojoVecError, match="ids"):
        sq.distances(q[:1], np.array([[20000]]))
