"""GPU parity at the shapes of the measured workloads.

- IVF-PQ 1M x 128 layout (d = 128, M = 16): the mixed form's deep-pool
  stream scan with the SQ8 pre-rerank, against the oracle's exact search over
  the probed rows;
- SQ8 / FP8 / BF16 Flat at d = 768 (the 1M x 768 embedding layout), several
  64-query tiles, k = 100 with k_factor 2 (strided splits of the 64-query
  form, and queries the per-query check sends back to the 16-query form);
- IVF-PQ 10M x 96 layout (d = 96, M = 12) above 4096 lists, per-query nprobe
  in 4..128 and k in {1, 10, 100}, with the autotuned schedule and a captured
  graph;
- the serving loop over the d = 128 / M = 16 IVF-PQ and the d = 768 SQ8
  cosine layouts.
Sizes are cut so the oracle finishes in seconds; the shapes (d, M, nlist
regime, k, k_factor) are the workloads'.
"""
import math

import numpy as np
import pytest

from oracle import oracle as o

pytestmark = pytest.mark.gpu


def clustered(seed, n, d, nclusters=50, sigma=12.0):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0, 128, (nclusters, d))
    lab = rng.integers(0, nclusters, n)
    x = centres[lab] + rng.normal(0, sigma, (n, d))
    return np.rint(np.clip(x, 0, 255)).astype(np.float32)


def unit_mixture(seed, n, d, ncentres=40, spread=0.3):
    """Unit-norm rows around random directions (the embedding recipe)."""
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((ncentres, d))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    x = c[rng.integers(0, ncentres, n)] + (spread / math.sqrt(d)) * rng.standard_normal((n, d))
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    return x.astype(np.float32)


def assert_same(Dg, Ig, Do, Io):
    assert np.array_equal(Ig, Io), f"id mismatch at {np.argwhere(Ig != Io)[:5].tolist()}"
    assert np.array_equal(Dg, Do), f"distance mismatch at {np.argwhere(Dg != Do)[:5].tolist()}"


def recall(gt, got):
    return float(np.mean([len(set(a) & set(b)) / len(a) for a, b in zip(gt.tolist(), got.tolist())]))


def _pair(gpu, x, d, nlist, M, metric, niter=3):
    oi = o.IVFPQIndex(d, nlist, M, metric)
    oi.train(x, niter_coarse=niter, niter_pq=niter)
    coarse, pq = oi.codebooks()
    gi = gpu.IndexIVFPQ(d, nlist, M, metric)
    gi.set_codebooks(coarse, pq)
    return oi, gi


# ------------------------------------------------- IVF-PQ d = 128, M = 16
@pytest.fixture(scope="module")
def ivfpq128(gpu):
    x = clustered(101, 24000, 128, nclusters=120)
    x[0, :], x[1, :] = 0.0, 255.0  # full code range: SQ8 scale 1, codes = rows
    oi, gi = _pair(gpu, x, 128, 48, 16, 0)
    oi.add(x)
    gi.add(x)
    q = clustered(102, 150, 128, nclusters=120)
    return x, q, oi, gi


def test_headline_mixed_deep_pool_every_probed_entry_is_exact(ivfpq128):
    # Mixed form, deep pool (k * k_factor >= 513), every probed entry a
    # candidate: after the FP32 rerank (with or without the SQ8 pre-rerank,
    # exact here) the result is the exact search over the probed rows, which
    # the oracle computes with its plain probe set and W = 64 rerank.
    x, q, oi, gi = ivfpq128
    k = 10
    sizes = np.sort(gi.list_sizes())[::-1]
    # the most probes whose largest lists still fit the 4096-candidate limit
    nprobe = int(np.searchsorted(np.cumsum(sizes), 4090, side="right"))
    kf = int(math.ceil(sizes[:nprobe].sum() / k))
    assert nprobe >= 2 and 513 <= k * kf <= 4096
    gi.nprobe = nprobe
    Do, Io = oi.search(q, k, nprobe, k_factor=kf)
    gi.set_search_mode("mixed")
    try:
        for pre in (0, 8):
            gi.set_prefilter(pre)
            for pool in ("hbm", "lds"):
                gi.set_deep_select(0, pool)
                Dg, Ig = gi.search(q, k, k_factor=kf)
                assert_same(Dg, Ig, Do, Io)
    finally:
        gi.set_prefilter(8)
        gi.set_deep_select(0, "hbm")
        gi.set_search_mode("plain")


def test_headline_mixed_deep_pool_matches_plain_at_kk_2880(ivfpq128):
    # The measured setting's pool width (k = 10, k_factor 288) over fewer
    # probed entries than the workload: the mixed deep-pool candidates,
    # narrowed by the pre-rerank, return the same top-10 as the plain form
    # reranked the same way (the mixed ADC distance differs from the plain
    # one only by FP32 rounding; the rerank distances are exact).
    x, q, oi, gi = ivfpq128
    gi.nprobe = 16
    gi.set_search_mode("plain")
    Dp, Ip = gi.search(q, 10, k_factor=288)
    Do, Io = oi.search(q, 10, 16, k_factor=288)
    assert_same(Dp, Ip, Do, Io)
    gi.set_search_mode("mixed")
    try:
        Dm, Im = gi.search(q, 10, k_factor=288)
    finally:
        gi.set_search_mode("plain")
    assert recall(Ip, Im) >= 0.995
    same = Ip == Im
    assert np.array_equal(Dm[same], Dp[same])


def test_headline_mixed_prefilter_on_float_rows_bounds(gpu):
    # Float rows (the codes only approximate them): the pre-rerank narrows
    # 2880 candidates to 80 by code distance. Stated bound: recall@10 against
    # the FP64 brute force drops by at most 0.01 against the FP32 rerank of all
    # 2880, and every returned distance is its row's FP32 distance, within
    # 1e-5 relative of FP64.
    rng = np.random.default_rng(111)
    x = (clustered(112, 24000, 128, nclusters=120) + rng.uniform(-0.5, 0.5, (24000, 128))).astype(np.float32)
    q = (clustered(113, 150, 128, nclusters=120) + rng.uniform(-0.5, 0.5, (150, 128))).astype(np.float32)
    _, gi = _pair(gpu, x, 128, 48, 16, 0)
    gi.add(x)
    gi.nprobe = 16
    gi.set_search_mode("mixed")
    try:
        gi.set_prefilter(0)
        Df, If = gi.search(q, 10, k_factor=288)
        gi.set_prefilter(8)
        Dp, Ip = gi.search(q, 10, k_factor=288)
    finally:
        gi.set_search_mode("plain")
    x64, q64 = x.astype(np.float64), q.astype(np.float64)
    d2 = (q64 ** 2).sum(1)[:, None] + (x64 ** 2).sum(1)[None] - 2 * q64 @ x64.T
    gt = np.argsort(d2, axis=1, kind="stable")[:, :10]
    assert recall(gt, If) - recall(gt, Ip) <= 0.01
    exact = np.take_along_axis(d2, Ip, 1)
    assert np.max(np.abs(Dp - exact) / np.maximum(exact, 1.0)) <= 1e-5


def _headline_pool(gi, k=10):
    """The most probes whose largest lists still fit the 4096-candidate limit,
    and the k_factor that makes every probed entry a candidate."""
    sizes = np.sort(gi.list_sizes())[::-1]
    nprobe = int(np.searchsorted(np.cumsum(sizes), 4090, side="right"))
    kf = int(math.ceil(sizes[:nprobe].sum() / k))
    assert nprobe >= 2 and 513 <= k * kf <= 4096
    return nprobe, kf


@pytest.mark.parametrize("pre", [0, 8])
def test_8bit_tables_every_probed_entry_is_exact(ivfpq128, pre):
    # 8-bit tables in the deep-pool scan only choose the candidates: with
    # every probed entry a candidate, the rerank returns the exact search over
    # the probed rows, as with FP32 tables.
    x, q, oi, gi = ivfpq128
    nprobe, kf = _headline_pool(gi)
    gi.nprobe = nprobe
    Do, Io = oi.search(q, 10, nprobe, k_factor=kf)
    gi.set_search_mode("mixed")
    gi.set_table_bits(8)
    gi.set_prefilter(pre)
    try:
        Dg, Ig = gi.search(q, 10, k_factor=kf)
    finally:
        gi.set_table_bits(32)
        gi.set_prefilter(8)
        gi.set_search_mode("plain")
    assert_same(Dg, Ig, Do, Io)


def test_8bit_tables_at_kk_2880_match_fp32_tables(ivfpq128):
    # The measured pool width over a third of the entries: the 8-bit tables'
    # candidates give the FP32 tables' top-10 (recall >= 0.99 between the
    # two), and a row both return carries the same exact rerank distance.
    x, q, oi, gi = ivfpq128
    gi.nprobe = 16
    gi.set_search_mode("mixed")
    try:
        D32, I32 = gi.search(q, 10, k_factor=288)
        gi.set_table_bits(8)
        D8, I8 = gi.search(q, 10, k_factor=288)
    finally:
        gi.set_table_bits(32)
        gi.set_search_mode("plain")
    assert recall(I32, I8) >= 0.99
    for i in range(q.shape[0]):
        d32 = dict(zip(I32[i].tolist(), D32[i].tolist()))
        for j, dv in zip(I8[i].tolist(), D8[i].tolist()):
            if j in d32:
                assert d32[j] == dv


@pytest.mark.parametrize("k,kf", [(600, 1), (10, 40)])
def test_8bit_tables_leave_other_shapes_on_fp32_tables(ivfpq128, k, kf):
    # No rerank (k_factor 1, the scan's distances are returned) or a pool
    # under the deep-pool threshold: the FP32 tables run, results unchanged.
    x, q, oi, gi = ivfpq128
    gi.nprobe = 16
    gi.set_search_mode("mixed")
    try:
        D32, I32 = gi.search(q, k, k_factor=kf)
        gi.set_table_bits(8)
        D8, I8 = gi.search(q, k, k_factor=kf)
    finally:
        gi.set_table_bits(32)
        gi.set_search_mode("plain")
    assert_same(D8, I8, D32, I32)


# ------------------------------------------------------- Flat at d = 768
@pytest.fixture(scope="module")
def emb768():
    x = unit_mixture(201, 12000, 768)
    q = unit_mixture(202, 300, 768)
    return x, q


@pytest.mark.parametrize("metric", [0, 1])
def test_sq8_d768_matches_oracle(gpu, emb768, metric):
    # cosine (int8 codes, symmetric) and L2 (affine uint8 rebased to int8) on
    # int8 MFMA at d = 768 over 300 queries (five 64-query tiles), k = 10 with
    # k_factor 1/2/4 and k = 100 with k_factor 2 (the 64-query form on nine
    # strided splits), with and without exclusions; then 1 query (the
    # 16-query form with its 768-key append buffer) and 17
    x, q = emb768
    gx = gpu.IndexFlatSQ8(768, metric)
    ox = o.SQ8Index(768, metric)
    for part in (x[:4000] * 0.5, x[4000:]):  # the second add widens the range
        gx.add(part)
        ox.add(part)
    assert gx.params() == ox.params()
    cg, ng = gx.codes()
    co, no = ox.codes()
    assert np.array_equal(cg, co) and np.array_equal(ng, no)
    ex = (np.random.default_rng(203).random(12000) < 0.2).astype(np.uint8)
    for k, kf in [(10, 1), (10, 2), (10, 4), (100, 2)]:
        Dg, Ig = gx.search(q, k, k_factor=kf)
        Do, Io = ox.search(q, k, k_factor=kf)
        assert_same(Dg, Ig, Do, Io)
    for k, kf in [(10, 2), (100, 2)]:
        Dg, Ig = gx.search(q, k, k_factor=kf, exclude=ex)
        Do, Io = ox.search(q, k, k_factor=kf, exclude=ex)
        assert_same(Dg, Ig, Do, Io)
    for nq in (1, 17):
        Dg, Ig = gx.search(q[:nq], 100, k_factor=2)
        Do, Io = ox.search(q[:nq], 100, k_factor=2)
        assert_same(Dg, Ig, Do, Io)


def test_sq8_large_append_buffer_small_rows(gpu):
    # d = 48 with k * k_factor = 256: the 16-query form's layout takes the
    # 768-key append buffer (flush threshold 512 appended keys)
    rng = np.random.default_rng(210)
    x = rng.integers(0, 9, (20000, 48)).astype(np.float32)
    q = rng.integers(0, 9, (300, 48)).astype(np.float32)
    ex = (rng.random(20000) < 0.25).astype(np.uint8)
    for metric in (0, 1):
        gx = gpu.IndexFlatSQ8(48, metric)
        ox = o.SQ8Index(48, metric)
        gx.add(x)
        ox.add(x)
        for nq in (1, 17, 300):
            for exm in (None, ex):
                Dg, Ig = gx.search(q[:nq], 128, k_factor=2, exclude=exm)
                Do, Io = ox.search(q[:nq], 128, k_factor=2, exclude=exm)
                assert_same(Dg, Ig, Do, Io)


def _period_layout(seed, n, d, nq, period):
    """Unit rows where every period-th row (ids 0, period, 2 period, ...) sits
    near one direction and the rest around 40 others; the first half of the
    queries sit at that direction, the rest among the other rows. Above 64
    candidates the 64-query form first bounds each query's keys by a sample
    pass over the rows 0, 16, 32, ..., then splits the rows by residue mod P
    (P = 10 for k x k_factor = 100, 13 for 200) under that bound. Period 130
    (both P divide it): a near query's nearest rows all fall in one split,
    more than the 64 its list keeps. Period 8: half of them lie in the
    sample, so its bound (the sample's 24th / 32nd key) is about that query's
    48th / 64th nearest row and fewer than k x k_factor keys pass. Either way
    the per-query check rejects the union and the 16-query form searches that
    query again."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(d)
    a /= np.linalg.norm(a)
    x = unit_mixture(seed + 1, n, d)
    near = a + (0.2 / math.sqrt(d)) * rng.standard_normal((len(range(0, n, period)), d))
    x[::period] = near / np.linalg.norm(near, axis=1, keepdims=True)
    q = unit_mixture(seed + 2, nq, d)
    qa = a + (0.2 / math.sqrt(d)) * rng.standard_normal((nq // 2, d))
    q[: nq // 2] = qa / np.linalg.norm(qa, axis=1, keepdims=True)
    return x.astype(np.float32), q.astype(np.float32)


@pytest.mark.parametrize("period", [130, 8])
@pytest.mark.parametrize("d", [768, 48])
@pytest.mark.parametrize("metric", [0, 1])
def test_sq8_strided_splits_reject_and_search_again(gpu, d, metric, period):
    # d = 768: the register-tile form, d = 48: the LDS form; 40 queries (three
    # 16-query tiles for the second search, the middle one mixed), k x k_factor
    # 200 and 100, with and without exclusions: equal to the oracle
    n = 14040  # 108 rows at multiples of 130
    x, q = _period_layout(250 + metric, n, d, 40, period)
    ex = (np.random.default_rng(252).random(n) < 0.2).astype(np.uint8)
    gx = gpu.IndexFlatSQ8(d, metric)
    ox = o.SQ8Index(d, metric)
    gx.add(x)
    ox.add(x)
    for k, kf in [(100, 2), (50, 2), (100, 1)]:
        for exm in (None, ex):
            Dg, Ig = gx.search(q, k, k_factor=kf, exclude=exm)
            Do, Io = ox.search(q, k, k_factor=kf, exclude=exm)
            assert_same(Dg, Ig, Do, Io)


def _w64_topk(x, q, k, metric):
    """Canonical top-k over every row of the rerank's exact distances
    (l2_distance_simd[64] / -inner_product_simd[64], the oracle's
    restatement), un-negated for inner products."""
    D = np.empty((q.shape[0], k), np.float32)
    I = np.empty((q.shape[0], k), np.int64)
    ids = np.arange(x.shape[0], dtype=np.int64)
    for i in range(q.shape[0]):
        y = np.ascontiguousarray(np.broadcast_to(q[i], x.shape))
        dd = o.l2_simd_batch(y, x, 64) if metric == 0 else -o.ip_simd_batch(y, x, 64)
        order = np.lexsort((ids, dd))[:k]
        D[i] = dd[order] if metric == 0 else -dd[order]
        I[i] = order
    return D, I


@pytest.mark.parametrize("form", ["mixed-fp8", "mixed-bf16"])
@pytest.mark.parametrize("metric", [0, 1])
def test_mixed_flat_d768_full_rerank_is_exact(gpu, emb768, form, metric):
    # FP8 E4M3 / BF16 candidates at d = 768 with every row reranked in FP32:
    # the exact W = 64 search, bit for bit; k_factor 2 keeps recall >= 0.95 of
    # it and its distances for the ids both return
    x, q = emb768
    x, q = x[:500], q[:130]
    ix = gpu.IndexFlat(768, metric)
    ix.add(x)
    Do, Io = _w64_topk(x, q, 10, metric)
    ix.set_search_mode(form, k_factor=50)
    Dg, Ig = ix.search(q, 10)
    assert_same(Dg, Ig, Do, Io)
    ix.set_search_mode(form, k_factor=2)
    D2, I2 = ix.search(q, 10)
    assert recall(Io, I2) >= 0.95
    same = I2 == Io
    assert np.array_equal(D2[same], Do[same])


@pytest.mark.parametrize("form", ["mixed-fp8", "mixed-bf16"])
@pytest.mark.parametrize("metric", [0, 1])
def test_mixed_flat_d768_candidates_exact_on_small_integers(gpu, form, metric):
    # Integers in [-15, 15] are exact in BF16 and, under the power-of-two row
    # scaling, in FP8 E4M3, and every product sum at d = 768 is exact in FP32:
    # the MFMA candidate distances equal the plain search over 300 queries
    # (several query tiles), ties and exclusions included.
    rng = np.random.default_rng(240 + metric)
    lo = -15 if metric == 1 else 0
    x = rng.integers(lo, 16, (8000, 768)).astype(np.float32)
    q = rng.integers(lo, 16, (300, 768)).astype(np.float32)
    ex = (rng.random(8000) < 0.2).astype(np.uint8)
    ix = gpu.IndexFlat(768, metric)
    ix.add(x)
    for k in (1, 10, 64, 100):  # 100: the 64-query form on strided splits
        ix.set_search_mode("plain")
        Dp, Ip = ix.search(q, k, exclude=ex)
        ix.set_search_mode(form)
        Dm, Im = ix.search(q, k, exclude=ex)
        assert_same(Dm, Im, Dp, Ip)
    Do, Io = o.flat_search(x, q[:40], 10, metric, exclude=ex)
    assert np.array_equal(ix.search(q[:40], 10, exclude=ex)[1], Io)


# ------------------------------------------- IVF-PQ d = 96, M = 12, > 4096 lists
@pytest.fixture(scope="module")
def ivfpq96(gpu):
    x = clustered(301, 36000, 96, nclusters=400)
    gi = gpu.IndexIVFPQ(96, 4160, 12)
    gi.train(x, niter_coarse=2, niter_pq=2)
    coarse, pq = gi.codebooks()
    oi = o.IVFPQIndex(96, 4160, 12)
    oi.set_codebooks(coarse, pq)
    oi.add(x)
    gi.add(x)
    q = clustered(302, 240, 96, nclusters=400)
    rng = np.random.default_rng(303)
    npq = rng.integers(4, 129, q.shape[0]).astype(np.int32)
    return x, q, npq, oi, gi


def test_ivfpq_d96_m12_lists_match_oracle(ivfpq96):
    x, q, npq, oi, gi = ivfpq96
    sizes = gi.list_sizes()
    for l in range(0, 4160, 97):
        ids_o, codes_o = oi.list(l)
        ids_g, codes_g = gi.get_list(l)
        assert sizes[l] == ids_o.size
        assert np.array_equal(ids_g, ids_o) and np.array_equal(codes_g, codes_o)
    assert sizes.sum() == x.shape[0]


@pytest.mark.parametrize("k", [1, 10, 100])
@pytest.mark.parametrize("kf", [1, 2])
def test_ivfpq_d96_m12_per_query_nprobe_matches_oracle(ivfpq96, k, kf):
    x, q, npq, oi, gi = ivfpq96
    gi.set_work_queue(0)
    Dg, Ig = gi.search(q, k, k_factor=kf, nprobe_per_query=npq)
    Do, Io = oi.search(q, k, 1, k_factor=kf, nprobe_per_query=npq)
    assert_same(Dg, Ig, Do, Io)
    gi.set_work_queue(1024)  # the skewed-list work queue: same results
    try:
        Dw, Iw = gi.search(q, k, k_factor=kf, nprobe_per_query=npq)
    finally:
        gi.set_work_queue(0)
    assert_same(Dw, Iw, Do, Io)


def test_ivfpq_d96_m12_autotune_and_graph_replay(ivfpq96):
    import torch

    x, q, npq, oi, gi = ivfpq96
    dev = torch.device("cuda:0")
    nq, k = q.shape[0], 10
    qd = torch.from_numpy(q).to(dev)
    npd = torch.from_numpy(npq).to(dev)
    D = torch.empty((nq, k), dtype=torch.float32, device=dev)
    I = torch.empty((nq, k), dtype=torch.int64, device=dev)
    name, ms = gi.autotune(qd.data_ptr(), nq, k, D.data_ptr(), I.data_ptr(), k_factor=2,
                           nprobe_ptr=npd.data_ptr(), max_nprobe=128)
    assert name in gi.SCHEDULES and ms > 0
    Do, Io = oi.search(q, k, 1, k_factor=2, nprobe_per_query=npq)
    De, Ie = gi.search(q, k, k_factor=2, nprobe_per_query=npq)  # the tuned schedule
    assert_same(De, Ie, Do, Io)
    g = gi.capture(qd.data_ptr(), nq, k, D.data_ptr(), I.data_ptr(), k_factor=2,
                   nprobe_ptr=npd.data_ptr(), max_nprobe=128)
    rng = np.random.default_rng(304)
    for rep in range(3):
        qn = clustered(305 + rep, nq, 96, nclusters=400)
        qd.copy_(torch.from_numpy(qn))
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        Do, Io = oi.search(qn, k, 1, k_factor=2, nprobe_per_query=npq)
        assert_same(D.cpu().numpy(), I.cpu().numpy(), Do, Io)
        qn2 = qn[rng.permutation(nq)]  # eager work between replays
        gi.search(qn2[:7], 3)
    del g
    gi.set_work_queue(0)
    gi.set_search_mode("plain", split_scan=False)


def test_ivfpq_d96_m12_mixed_mfma_coarse_bound(ivfpq96):
    # 4160 lists: the mixed form's coarse step is the bf16 GEMM form; with
    # every probed entry reranked the results stay within the probe-set
    # difference of near-tie centroids
    x, q, npq, oi, gi = ivfpq96
    gi.set_search_mode("mixed")
    try:
        Dm, Im = gi.search(q, 10, k_factor=40, nprobe_per_query=npq)
    finally:
        gi.set_search_mode("plain")
    Do, Io = oi.search(q, 10, 1, k_factor=40, nprobe_per_query=npq)
    assert recall(Io, Im) >= 0.97
    same = Im == Io
    assert np.array_equal(Dm[same], Do[same])


# -------------------------------------------------------------- serving
def test_server_over_ivfpq_128_m16_mixed(gpu, ivfpq128):
    # the serving loop over the d = 128 / M = 16 layout in the mixed form
    # with the pre-rerank: every request equals the same request alone
    x, q, oi, gi = ivfpq128
    gi.set_search_mode("mixed")
    gi.set_prefilter(8)
    rng = np.random.default_rng(120)
    srv = gpu.Server(gi, slots=4, max_batch=512, mem_budget=256 << 20)
    try:
        reqs = []
        for i in range(30):
            nq = int(rng.integers(1, 200))
            qq = clustered(130 + i, nq, 128, nclusters=120)
            k = int(rng.choice([10, 100]))
            nprobe = int(rng.integers(8, 48))
            kf = int(rng.choice([1, 4, 40]))
            reqs.append((qq, k, nprobe, kf, srv.submit(qq, k=k, nprobe=nprobe, k_factor=kf)))
        outs = [(qq, k, nprobe, kf, srv.wait(t)) for qq, k, nprobe, kf, t in reqs]
        for qq, k, nprobe, kf, (D, I, lat) in outs:
            gi.nprobe = nprobe
            Da, Ia = gi.search(qq, k, k_factor=kf)
            assert_same(D, I, Da, Ia)
    finally:
        srv.close()
        gi.set_search_mode("plain")


def test_server_over_sq8_d768_cosine(gpu, emb768):
    x, q = emb768
    ix = gpu.IndexFlatSQ8(768, 1)
    ix.add(x)
    rng = np.random.default_rng(220)
    srv = gpu.Server(ix, slots=3, max_batch=400, mem_budget=512 << 20)
    try:
        reqs = []
        for i in range(16):
            nq = int(rng.integers(1, 150))
            qq = unit_mixture(230 + i, nq, 768)
            k = int(rng.choice([10, 100]))
            kf = 2
            reqs.append((qq, k, kf, srv.submit(qq, k=k, nprobe=1, k_factor=kf)))
        outs = [(qq, k, kf, srv.wait(t)) for qq, k, kf, t in reqs]
        for qq, k, kf, (D, I, lat) in outs:
            Da, Ia = ix.search(qq, k, k_factor=kf)
            assert_same(D, I, Da, Ia)
    finally:
        srv.close()


def test_one_loop_serves_ivfpq_128_and_sq8_768(gpu, ivfpq128, emb768):
    # both layouts resident and served by one dispatcher, one set of slots and
    # one memory budget; requests of the two indexes arrive interleaved, with
    # mixed k, and each result equals the same request searched alone
    x, q, oi, gi = ivfpq128
    xe, qe = emb768
    sx = gpu.IndexFlatSQ8(768, 1)
    sx.add(xe)
    gi.set_search_mode("mixed")
    rng = np.random.default_rng(250)
    srv = gpu.Server.both(gi, sx, slots=4, max_batch=400, mem_budget=512 << 20)
    try:
        reqs = []
        for i in range(40):
            k = int(rng.choice([10, 100]))
            nq = int(rng.integers(1, 120))
            if i % 2 == 0:
                qq = clustered(260 + i, nq, 128, nclusters=120)
                nprobe, kf = int(rng.integers(8, 48)), int(rng.choice([1, 4, 30]))
                reqs.append(("ivf", qq, k, nprobe, kf, srv.submit(qq, k=k, nprobe=nprobe, k_factor=kf)))
            else:
                qq = unit_mixture(300 + i, nq, 768)
                reqs.append(("sq8", qq, k, 1, 2, srv.submit(qq, k=k, nprobe=1, k_factor=2, target=sx)))
        outs = [(kind, qq, k, npb, kf, srv.wait(t)) for kind, qq, k, npb, kf, t in reqs]
        st = srv.stats()
        assert st["completed_queries"] == sum(r[1].shape[0] for r in reqs)
        for kind, qq, k, npb, kf, (D, I, lat) in outs:
            if kind == "ivf":
                gi.nprobe = npb
                Da, Ia = gi.search(qq, k, k_factor=kf)
            else:
                Da, Ia = sx.search(qq, k, k_factor=kf)
            assert_same(D, I, Da, Ia)
    finally:
        srv.close()
        gi.set_search_mode("plain")
