"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle.

Bar: bit-exact ids, SQ8 codes and integer dots; FP32 distances bit-exact as
well wherever the plain form mirrors the oracle's operation order (asserted
with equality; the stated tolerance of 1e-5 relative is the fallback bound).
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


def clustered(seed, n, d, nclusters=50, sigma=12.0):
    """Integer-valued Gaussian mixture in [0, 255] (the Flat / IVF-PQ data recipe)."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0, 128, (nclusters, d))
    lab = rng.integers(0, nclusters, n)
    x = centres[lab] + rng.normal(0, sigma, (n, d))
    return np.rint(np.clip(x, 0, 255)).astype(np.float32)


def assert_same(Dg, Ig, Do, Io):
    assert np.array_equal(Ig, Io), f"id mismatch at {np.argwhere(Ig != Io)[:5].tolist()}"
    assert np.array_equal(Dg, Do) or np.allclose(Dg, Do, rtol=RTOL, atol=0)


# ------------------------------------------------------------------ Flat
@pytest.mark.parametrize("d", [128, 96, 64, 32, 17, 8, 4])
@pytest.mark.parametrize("metric", [0, 1])
def test_flat_matches_oracle(gpu, d, metric):
    rng = np.random.default_rng(d + 10 * metric)
    base = rng.standard_normal((3000, d)).astype(np.float32)
    q = rng.standard_normal((37, d)).astype(np.float32)
    ix = gpu.IndexFlat(d, metric)
    ix.add(base[:1000])
    ix.add(base[1000:])
    for k in (1, 10, 100):
        Dg, Ig = ix.search(q, k)
        Do, Io = o.flat_search(base, q, k, metric)
        assert_same(Dg, Ig, Do, Io)


def test_flat_integer_clustered_bit_exact(gpu):
    base = clustered(1, 20000, 128)
    q = clustered(2, 300, 128)
    ix = gpu.IndexFlat(128)
    ix.add(base)
    Dg, Ig = ix.search(q, 10)
    Do, Io = o.flat_search(base, q, 10)
    assert np.array_equal(Ig, Io) and np.array_equal(Dg, Do)


@pytest.mark.parametrize("metric", [0, 1])
def test_flat_mixed_bf16_mfma_is_exact_on_integer_data(gpu, metric):
    # Integers up to 256 are exact in bf16 and every partial sum stays below
    # 2^24, so the GEMM-form distance on MFMA equals the FP32 result.
    base = clustered(11, 9000, 128)
    q = clustered(12, 100, 128)
    if metric == 1:
        base, q = base - 128, q - 128  # signed values keep IP ranking non-trivial
    ix = gpu.IndexFlat(128, metric)
    ix.add(base)
    Dp, Ip = ix.search(q, 10)
    ix.set_search_mode("mixed")
    Dm, Im = ix.search(q, 10)
    assert np.array_equal(Im, Ip) and np.array_equal(Dm, Dp)
    ix.set_search_mode("mixed", k_factor=4)
    Dr, Ir = ix.search(q, 10)
    assert np.array_equal(Ir, Ip) and np.array_equal(Dr, Dp)
    # k = 100: 100 candidates per query, the 64-query form's LDS layout on
    # strided splits under a sampled bound (exact here, so equal as well)
    ix.set_search_mode("plain")
    Dp, Ip = ix.search(q, 100)
    ix.set_search_mode("mixed")
    Dm, Im = ix.search(q, 100)
    assert np.array_equal(Im, Ip) and np.array_equal(Dm, Dp)


def test_flat_mixed_rerank_on_float_data(gpu):
    rng = np.random.default_rng(13)
    base = rng.standard_normal((20000, 96)).astype(np.float32)
    q = rng.standard_normal((64, 96)).astype(np.float32)
    ix = gpu.IndexFlat(96)
    ix.add(base)
    ix.set_search_mode("mixed", k_factor=8)
    Dm, Im = ix.search(q, 10)
    Do, Io = o.flat_search(base, q, 10)
    assert np.mean(Im == Io) > 0.99
    same = Im == Io
    assert np.allclose(Dm[same], Do[same], rtol=1e-5)


def test_flat_ties_exclusion_and_padding(gpu):
    g = np.load(os.path.join(GOLDEN, "ties.npz"))
    ix = gpu.IndexFlat(16)
    ix.add(g["base"])
    D, I = ix.search(g["q"], int(g["k"]))
    assert np.array_equal(I, g["ids"]) and np.array_equal(D, g["dist"])
    ex = np.zeros(300, np.uint8)
    ex[::3] = 1
    Dg, Ig = ix.search(g["q"], 25, exclude=ex)
    Do, Io = o.flat_search(g["base"], g["q"], 25, exclude=ex)
    assert_same(Dg, Ig, Do, Io)
    small = gpu.IndexFlat(4)
    small.add(np.ones((3, 4), np.float32))
    D, I = small.search(np.zeros((2, 4), np.float32), 5)
    assert I[:, 3:].tolist() == [[-1, -1]] * 2 and (D[:, 3:] == 0).all()


# ------------------------------------------------------------------ k-means / PQ
@pytest.mark.parametrize("d,k,n", [(128, 64, 5000), (8, 256, 3000), (4, 16, 500), (33, 40, 2000)])
def test_kmeans_matches_oracle(gpu, d, k, n):
    rng = np.random.default_rng(d * k)
    x = rng.standard_normal((n, d)).astype(np.float32) * 10
    init, off = o.kmeans_init(x, k)
    co, ao = o.kmeans_train(x, k, 6, init=init, sample_offset=off)
    km = gpu.KMeans(d, k, 6)
    cg = km.train(x, init=init, sample_offset=off)
    assert np.array_equal(km.assignments, ao)
    assert np.array_equal(cg, co)
    # seeded init path (no explicit init) agrees too
    cg2 = gpu.KMeans(d, k, 3).train(x)
    co2, _ = o.kmeans_train(x, k, 3)
    assert np.array_equal(cg2, co2)


def test_pq_train_encode_lut_match_oracle(gpu):
    rng = np.random.default_rng(21)
    x = rng.standard_normal((4000, 64)).astype(np.float32)
    pq = gpu.ProductQuantizer(64, 8, niter=4)
    pq.train(x)
    cbo = o.pq_train(x, 8, niter=4)
    assert np.array_equal(pq.centroids, cbo)
    codes = pq.compute_codes(x)
    assert np.array_equal(codes, o.pq_encode(x, cbo))
    q = rng.standard_normal((5, 64)).astype(np.float32)
    for metric in (0, 1):
        lut = pq.compute_distance_table(q, metric)
        for i in range(5):
            assert np.array_equal(lut[i], o.pq_lut(q[i], cbo, metric))


def test_pq_adc_identity(gpu):
    x = (np.random.default_rng(191).random((100, 16)) * 2 - 1).astype(np.float32)
    pq = gpu.ProductQuantizer(16, 4)
    pq.train(x)
    codes = pq.compute_codes(x)
    lut = pq.compute_distance_table(x[:1])[0]
    adc = np.float32(0)
    for m in range(4):
        adc = np.float32(adc + lut[m, codes[0, m]])
    dec = pq.decode(codes[:1])[0]
    assert abs(float(adc) - float(((x[0].astype(np.float64) - dec) ** 2).sum())) < 1e-5


# ------------------------------------------------------------------ IVF-PQ
def _pair(gpu, x, d, nlist, M, metric, niter=4):
    oi = o.IVFPQIndex(d, nlist, M, metric)
    oi.train(x, niter_coarse=niter, niter_pq=niter)
    coarse, pq = oi.codebooks()
    gi = gpu.IndexIVFPQ(d, nlist, M, metric)
    gi.set_codebooks(coarse, pq)
    return oi, gi


@pytest.mark.parametrize("metric", [0, 1])
def test_ivfpq_train_matches_oracle(gpu, metric):
    x = clustered(3, 6000, 32, nclusters=40)
    if metric == 1:
        x = x / 255.0
    gi = gpu.IndexIVFPQ(32, 32, 8, metric)
    gi.train(x, niter_coarse=5, niter_pq=5)
    oi = o.IVFPQIndex(32, 32, 8, metric)
    oi.train(x, niter_coarse=5, niter_pq=5)
    cg, pg = gi.codebooks()
    co, po = oi.codebooks()
    assert np.array_equal(cg, co)
    assert np.array_equal(pg, po)


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("metric", [0, 1])
@pytest.mark.parametrize("M", [16, 8, 4])
def test_ivfpq_add_and_search_match_oracle(gpu, metric, M, split):
    x = clustered(4, 8000, 64, nclusters=60)
    if metric == 1:
        x = x / 255.0
    oi, gi = _pair(gpu, x, 64, 40, M, metric)
    gi.set_search_mode("plain", split_scan=split)
    oi.add(x[:5000])
    oi.add(x[5000:])
    gi.add(x[:5000])
    gi.add(x[5000:])
    sizes = gi.list_sizes()
    for l in range(40):
        ids_o, codes_o = oi.list(l)
        ids_g, codes_g = gi.get_list(l)
        assert sizes[l] == ids_o.size
        assert np.array_equal(ids_g, ids_o) and np.array_equal(codes_g, codes_o)
    q = clustered(5, 50, 64, nclusters=60)
    if metric == 1:
        q = q / 255.0
    for nprobe, k, kf in [(1, 10, 1), (8, 10, 1), (40, 10, 1), (8, 100, 1), (8, 10, 4), (3, 1, 2)]:
        gi.nprobe = nprobe
        Dg, Ig = gi.search(q, k, k_factor=kf)
        Do, Io = oi.search(q, k, nprobe, k_factor=kf)
        assert_same(Dg, Ig, Do, Io)


@pytest.mark.parametrize("wq,metric", [(0, 0), (300, 0), (4096, 0), (0, 1)])
def test_ivfpq_per_query_nprobe_exclusion_padding(gpu, wq, metric):
    x = clustered(6, 4000, 32, nclusters=30)
    if metric == 1:
        x = x / 255.0
    oi, gi = _pair(gpu, x, 32, 64, 8, metric)
    gi.set_work_queue(wq)
    oi.add(x)
    gi.add(x)
    q = clustered(7, 40, 32, nclusters=30)
    if metric == 1:
        q = q / 255.0
    npq = np.random.default_rng(0).integers(1, 70, 40).astype(np.int32)
    ex = np.zeros(4000, np.uint8)
    ex[::7] = 1
    Dg, Ig = gi.search(q, 30, nprobe_per_query=npq, exclude=ex)
    Do, Io = oi.search(q, 30, 1, nprobe_per_query=npq, exclude=ex)
    assert_same(Dg, Ig, Do, Io)
    assert not np.isin(Ig[Ig >= 0], np.nonzero(ex)[0]).any()
    gi.nprobe = 1
    Dg, Ig = gi.search(q[:3], 2000)  # more than one list holds: padded tail
    Do, Io = oi.search(q[:3], 2000, 1)
    assert_same(Dg, Ig, Do, Io)
    assert (Ig[:, -1] == -1).all() and (Dg[:, -1] == np.float32(1e38)).all()


@pytest.mark.parametrize("mode", ["query", "split", "wq"])
def test_ivfpq_long_lists_cross_scan_chunks(gpu, mode):
    # one dominant cluster so a list exceeds the 4096-entry scan chunk
    rng = np.random.default_rng(8)
    x = np.rint(np.clip(rng.normal(100, 3, (20000, 16)), 0, 255)).astype(np.float32)
    oi, gi = _pair(gpu, x, 16, 4, 4, 0)
    gi.set_search_mode("plain", split_scan=mode == "split")
    gi.set_work_queue(2048 if mode == "wq" else 0)
    oi.add(x)
    gi.add(x)
    assert gi.list_sizes().max() > 4096
    q = np.rint(np.clip(rng.normal(100, 3, (20, 16)), 0, 255)).astype(np.float32)
    gi.nprobe = 4
    for k, kf in ((10, 1), (500, 1), (10, 3)):
        Dg, Ig = gi.search(q, k, k_factor=kf)
        Do, Io = oi.search(q, k, 4, k_factor=kf)
        assert_same(Dg, Ig, Do, Io)


@pytest.mark.parametrize("pool", ["lds", "hbm"])
@pytest.mark.parametrize("d,M,metric", [(128, 16, 0), (64, 8, 0), (32, 8, 1), (48, 6, 1)])
def test_ivfpq_deep_select_matches_oracle(gpu, d, M, metric, pool):
    # deep-pool scan (keys in HBM, one radix select per query) on every
    # k * k_factor: same canonical top-k as the oracle
    x = clustered(20 + d, 6000, d, nclusters=40)
    if metric == 1:
        x = x / 255.0
    oi, gi = _pair(gpu, x, d, 32, M, metric)
    oi.add(x)
    gi.add(x)
    gi.set_deep_select(1, pool)
    q = clustered(21 + d, 30, d, nclusters=40)
    if metric == 1:
        q = q / 255.0
    for nprobe, k, kf in [(4, 10, 1), (8, 700, 1), (32, 10, 60), (32, 1000, 1), (32, 3, 1)]:
        gi.nprobe = nprobe
        Dg, Ig = gi.search(q, k, k_factor=kf)
        Do, Io = oi.search(q, k, nprobe, k_factor=kf)
        assert_same(Dg, Ig, Do, Io)
    npq = np.random.default_rng(1).integers(1, 40, 30).astype(np.int32)
    ex = np.zeros(6000, np.uint8)
    ex[::5] = 1
    Dg, Ig = gi.search(q, 800, nprobe_per_query=npq, exclude=ex)
    Do, Io = oi.search(q, 800, 1, nprobe_per_query=npq, exclude=ex)
    assert_same(Dg, Ig, Do, Io)
    assert not np.isin(Ig[Ig >= 0], np.nonzero(ex)[0]).any()


@pytest.mark.parametrize("metric", [0, 1])
def test_ivfpq_deep_select_ties_break_by_lower_id(gpu, metric):
    # every row 12 times: runs of equal codes and equal ADC distances, so the
    # k-th key is shared by keys on both sides of the cut
    x = np.repeat(clustered(30, 400, 16, nclusters=8), 12, axis=0)
    if metric == 1:
        x = x / 255.0
    oi, gi = _pair(gpu, x, 16, 8, 4, metric)
    oi.add(x)
    gi.add(x)
    q = clustered(31, 20, 16, nclusters=8)
    if metric == 1:
        q = q / 255.0
    gi.nprobe = 8
    for k in (515, 1001, 7, 2500):
        Do, Io = oi.search(q, k, 8)
        # deep form with either pool, and the running top-k form
        for deep_min, pool in ((1, "lds"), (1, "hbm"), (100000, "lds")):
            gi.set_deep_select(deep_min, pool)
            Dg, Ig = gi.search(q, k)
            assert_same(Dg, Ig, Do, Io)


def test_ivfpq_deep_select_mixed_equals_running_form(gpu):
    x = clustered(32, 8000, 64, nclusters=40)
    oi, gi = _pair(gpu, x, 64, 32, 16, 0)
    gi.add(x)
    gi.set_search_mode("mixed")
    gi.nprobe = 12
    q = clustered(33, 40, 64, nclusters=40)
    for k, kf in ((600, 1), (10, 80), (3000, 1)):
        gi.set_deep_select(1, "lds")
        Dd, Id = gi.search(q, k, k_factor=kf)
        gi.set_deep_select(1, "hbm")
        Dh, Ih = gi.search(q, k, k_factor=kf)
        gi.set_deep_select(100000)
        Dr, Ir = gi.search(q, k, k_factor=kf)
        assert np.array_equal(Id, Ir) and np.array_equal(Dd, Dr)
        assert np.array_equal(Ih, Ir) and np.array_equal(Dh, Dr)
    gi.set_search_mode("plain")


def test_ivfpq_mixed_form_error_bound_and_rerank(gpu):
    # Mixed L2 tables: |q - c_l|^2 + T[l] + Q[q] equals the residual ADC
    # distance up to FP32 rounding; with every candidate reranked the result
    # is the exact search.
    x = clustered(9, 6000, 32, nclusters=30)
    oi, gi = _pair(gpu, x, 32, 16, 8, 0)
    oi.add(x)
    gi.add(x)
    q = clustered(10, 30, 32, nclusters=30)
    gi.nprobe = 16
    gi.set_search_mode("mixed")
    Dm, Im = gi.search(q, 50)
    Do, Io = oi.search(q, 50, 16)
    assert np.allclose(Dm, Do, rtol=1e-4, atol=0.5)
    overlap = np.mean([len(set(a) & set(b)) / 50 for a, b in zip(Im.tolist(), Io.tolist())])
    assert overlap > 0.95
    gi.set_prefilter(0)  # no SQ8 pre-rerank: all candidates in FP32
    Dr, Ir = gi.search(q[:5], 5, k_factor=800)  # 4000 of 6000 candidates reranked
    Df, If = o.flat_search(x, q[:5], 5)
    assert np.array_equal(Ir, If) and np.array_equal(Dr, Df)
    gi.set_search_mode("plain")


@pytest.mark.parametrize("metric", [0, 1])
def test_ivfpq_mixed_prefilter_exact_codes(gpu, metric):
    # Integer rows whose range pins the SQ8 scale to 1 (L2: [0, 255]; IP:
    # max |x| = 127): the codes are the rows, the SQ8 pre-rerank orders
    # candidates by the exact distance with the same id tie-break, so
    # narrowing k * k_factor candidates to max(64, 8k) before the FP32 rerank
    # changes nothing.
    x = clustered(60 + metric, 6000, 32, nclusters=30)
    if metric == 0:
        x[0, :], x[1, :] = 0.0, 255.0
    else:
        x = np.clip(x - 128.0, -127.0, 127.0)
        x[0, :] = 127.0
    _, gi = _pair(gpu, x, 32, 16, 8, metric)
    gi.add(x)
    q = clustered(70 + metric, 40, 32, nclusters=30)
    if metric == 1:
        q = np.clip(q - 128.0, -127.0, 127.0)
    gi.nprobe = 8
    gi.set_search_mode("mixed")
    for k, kf in ((10, 50), (100, 12), (1, 300)):
        gi.set_prefilter(0)
        Df, If = gi.search(q, k, k_factor=kf)
        gi.set_prefilter(8)
        Dp, Ip = gi.search(q, k, k_factor=kf)
        assert np.array_equal(Ip, If) and np.array_equal(Dp, Df), (k, kf)
    gi.set_search_mode("plain")


def test_ivfpq_mixed_prefilter_float_rows(gpu):
    # Rows the codes only approximate: the narrowed candidates still hold
    # nearly all of the FP32 top-k, and every returned distance is the FP32
    # rerank distance of its row.
    rng = np.random.default_rng(80)
    x = (clustered(81, 6000, 32, nclusters=30) + rng.uniform(-0.5, 0.5, (6000, 32))).astype(np.float32)
    _, gi = _pair(gpu, x, 32, 16, 8, 0)
    gi.add(x)
    q = clustered(82, 50, 32, nclusters=30)
    gi.nprobe = 8
    gi.set_search_mode("mixed")
    gi.set_prefilter(0)
    Df, If = gi.search(q, 10, k_factor=60)
    gi.set_prefilter(8)
    Dp, Ip = gi.search(q, 10, k_factor=60)
    overlap = np.mean([len(set(a) & set(b)) / 10 for a, b in zip(Ip.tolist(), If.tolist())])
    assert overlap > 0.99
    for i in range(q.shape[0]):  # a row kept by both has the same FP32 rerank distance
        full = dict(zip(If[i].tolist(), Df[i].tolist()))
        for j, dv in zip(Ip[i].tolist(), Dp[i].tolist()):
            if j in full:
                assert full[j] == dv
    Da, Ia = o.flat_search(x, q, 6000)  # every row's distance: within FP32 rounding
    for i in range(q.shape[0]):
        exact = dict(zip(Ia[i].tolist(), Da[i].tolist()))
        assert np.allclose([exact[j] for j in Ip[i].tolist()], Dp[i], rtol=1e-6)
    gi.set_search_mode("plain")


def test_ivfpq_mixed_mfma_coarse_large_nlist(gpu):
    # above 4096 lists the mixed form takes its coarse top-nprobe from the
    # bf16 GEMM-form distance on MFMA: the probe sets may differ from the FP32
    # ones only at near-ties, so results stay within the mixed-form bound
    x = clustered(40, 9000, 16, nclusters=60)
    oi, gi = _pair(gpu, x, 16, 4200, 4, 0, niter=1)
    oi.add(x)
    gi.add(x)
    q = clustered(41, 40, 16, nclusters=60)
    gi.nprobe = 64
    gi.set_search_mode("mixed")
    Dm, Im = gi.search(q, 10)
    Do, Io = oi.search(q, 10, 64)
    overlap = np.mean([len(set(a) & set(b)) / 10 for a, b in zip(Im.tolist(), Io.tolist())])
    assert overlap > 0.9
    assert np.allclose(np.sort(Dm, 1)[:, :5], np.sort(Do, 1)[:, :5], rtol=1e-3, atol=1.0)
    # 100 probes: the coarse top-100 on strided splits of the centroids
    gi.nprobe = 100
    Dm, Im = gi.search(q, 10)
    Do, Io = oi.search(q, 10, 100)
    overlap = np.mean([len(set(a) & set(b)) / 10 for a, b in zip(Im.tolist(), Io.tolist())])
    assert overlap > 0.9
    gi.set_search_mode("plain")


def test_ivfpq_untrained_and_empty(gpu):
    gi = gpu.IndexIVFPQ(8, 4, 2)
    D, I = gi.search(np.zeros((2, 8), np.float32), 3)
    assert (I == -1).all() and (D == np.float32(1e38)).all()
    gi.add(np.ones((10, 8), np.float32))  # untrained add is silent
    assert gi.ntotal == 0


# ------------------------------------------------------------------ SQ8
@pytest.mark.parametrize("use_mfma", [True, False])
@pytest.mark.parametrize("d", [128, 100, 3])
@pytest.mark.parametrize("metric", [0, 1])
def test_sq8_codes_and_search_match_oracle(gpu, metric, d, use_mfma):
    rng = np.random.default_rng(30 + metric + d)
    x = rng.standard_normal((6000, d)).astype(np.float32)
    gx = gpu.IndexFlatSQ8(d, metric)
    gx.set_search_mode(use_mfma)
    ox = o.SQ8Index(d, metric)
    for part in (x[:2000] * 0.5, x[2000:]):  # second batch widens the range
        gx.add(part)
        ox.add(part)
    assert gx.params() == ox.params()
    cg, ng = gx.codes()
    co, no = ox.codes()
    assert np.array_equal(cg, co) and np.array_equal(ng, no)
    q = rng.standard_normal((33, d)).astype(np.float32)
    for k, kf in [(1, 1), (10, 1), (10, 2), (10, 4), (100, 1)]:
        Dg, Ig = gx.search(q, k, k_factor=kf)
        Do, Io = ox.search(q, k, k_factor=kf)
        assert_same(Dg, Ig, Do, Io)


@pytest.mark.parametrize("metric", [0, 1])
def test_sq8_many_query_tiles_with_exclusion(gpu, metric):
    # 300 queries span several 64-query workgroup tiles of the int8 MFMA form
    # (last one partial); codes of a few distinct values give many exact ties.
    rng = np.random.default_rng(60 + metric)
    x = rng.integers(0, 6, (15000, 48)).astype(np.float32)
    q = rng.integers(0, 6, (300, 48)).astype(np.float32)
    ex = (rng.random(15000) < 0.25).astype(np.uint8)
    gx = gpu.IndexFlatSQ8(48, metric)
    ox = o.SQ8Index(48, metric)
    gx.add(x)
    ox.add(x)
    # (100, 2), (50, 3): above 64 candidates the 64-query form on strided
    # splits (one top 64 per split); 1 query: the 16-query form, its split
    # count from the CU balance and flushes by radix select
    for k, kf in [(1, 1), (10, 1), (16, 4), (64, 1), (10, 6), (100, 2), (50, 3)]:
        Dg, Ig = gx.search(q, k, k_factor=kf, exclude=ex)
        Do, Io = ox.search(q, k, k_factor=kf, exclude=ex)
        assert_same(Dg, Ig, Do, Io)
    for nq in (1, 17):  # one tile: the most splits
        Dg, Ig = gx.search(q[:nq], 100, k_factor=2, exclude=ex)
        Do, Io = ox.search(q[:nq], 100, k_factor=2, exclude=ex)
        assert_same(Dg, Ig, Do, Io)


def test_sq8_ip_ordering_and_scale(gpu):
    ix = gpu.IndexFlatSQ8(4, 1)
    ix.add(np.array([[1, 0, 0, 0], [-1, 0, 0, 0], [0.5, 0, 0, 0]], np.float32))
    D, I = ix.search(np.array([[1, 0, 0, 0]], np.float32), 3)
    assert I[0].tolist() == [0, 2, 1] and D[0, 0] > D[0, 1] > D[0, 2]
    ix = gpu.IndexFlatSQ8(4, 1)
    ix.add(np.array([[1.0, 0.25, 0, 0]], np.float32))
    ix.add(np.array([[-4.0, 0.0, 0, 0]], np.float32))
    assert abs(ix.scale - 4.0 / 127.0) < 1e-6
