"""GPU parity of the mixed IVF-PQ scan on FP16 MFMA tiles (table_bits 16).

The tile scan computes |q - c_l|^2 + pre[e] - 2 <q, r_e> with the query tile
and the PQ-decoded entries on v_mfma_f32_16x16x32_f16 and keeps, per query,
exactly the k * k_factor smallest (distance, id) pairs of those distances for
the exact rerank.

Anchor: with integer data, integer coarse centroids and small-integer PQ
codebooks every FP16 operand is exact, every product and sum is an integer
below 2^24, and the tile distance equals the reference's residual ADC distance
(index_ivf_pq.mojo:316-347) exactly. The candidate sets are then the plain
form's, so after the rerank the results equal the oracle's bit for bit, ties
by id included -- which pins the histogram windows, the boundary select and
the fallback of the scan. On float data the tile distances are approximations
that only choose candidates; there the bar is the exact search over every
probed entry and the FP32-table form's recall.
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


def assert_same(Dg, Ig, Do, Io):
    assert np.array_equal(Ig, Io), f"id mismatch at {np.argwhere(Ig != Io)[:5].tolist()}"
    assert np.array_equal(Dg, Do), f"distance mismatch at {np.argwhere(Dg != Do)[:5].tolist()}"


def recall(gt, got):
    return float(np.mean([len(set(a) & set(b)) / len(a) for a, b in zip(gt.tolist(), got.tolist())]))


def int_index(gpu, x, d, nlist, M, seed):
    """Oracle and GPU indexes sharing integer coarse centroids (rows of x)
    and PQ codebooks with entries in [-6, 6]."""
    rng = np.random.default_rng(seed)
    coarse = x[rng.choice(x.shape[0], nlist, replace=False)].copy()
    pq = rng.integers(-6, 7, (M, 256, d // M)).astype(np.float32)
    oi = o.IVFPQIndex(d, nlist, M, 0)
    oi.set_codebooks(coarse, pq)
    gi = gpu.IndexIVFPQ(d, nlist, M, 0)
    gi.set_codebooks(coarse, pq)
    return oi, gi


def tile_search(gi, q, k, kf, **kw):
    # these batches are smaller than the tile minimum (16 queries per CU):
    # a minimum of 1 sends every batch to the tile scan
    gi.set_search_mode("mixed")
    gi.set_table_bits(16)
    gi.set_tile_min_batch(1)
    try:
        return gi.search(q, k, k_factor=kf, **kw)
    finally:
        gi.set_tile_min_batch(0)
        gi.set_table_bits(32)
        gi.set_search_mode("plain")


@pytest.fixture(scope="module")
def int128(gpu):
    x = clustered(701, 20000, 128, nclusters=60)
    x[0, :], x[1, :] = 0.0, 255.0  # full range: SQ8 scale 1 (exact pre-rerank codes)
    oi, gi = int_index(gpu, x, 128, 40, 16, 702)
    oi.add(x)
    gi.add(x)
    q = clustered(703, 120, 128, nclusters=60)
    return x, q, oi, gi


@pytest.mark.parametrize("k,kf", [(10, 288), (10, 40), (100, 2), (1, 700), (5, 3)])
@pytest.mark.parametrize("pre", [0, 8])
def test_tile_scan_equals_oracle_on_exact_integers(int128, k, kf, pre):
    # kk from 15 to 2880 over ~4000 probed entries per query
    x, q, oi, gi = int128
    gi.nprobe = 8
    gi.set_prefilter(pre)
    try:
        Dg, Ig = tile_search(gi, q, k, kf)
    finally:
        gi.set_prefilter(8)
    Do, Io = oi.search(q, k, 8, k_factor=kf)
    assert_same(Dg, Ig, Do, Io)


def test_tile_scan_exclusion_and_per_query_nprobe(int128):
    x, q, oi, gi = int128
    rng = np.random.default_rng(710)
    excl = (rng.random(x.shape[0]) < 0.25).astype(np.uint8)
    npq = rng.integers(1, 13, q.shape[0]).astype(np.int32)
    gi.nprobe = 12
    for k, kf in ((10, 288), (20, 5)):
        Dg, Ig = tile_search(gi, q, k, kf, nprobe_per_query=npq, exclude=excl)
        Do, Io = oi.search(q, k, 12, k_factor=kf, nprobe_per_query=npq, exclude=excl)
        assert_same(Dg, Ig, Do, Io)


def test_tile_scan_single_and_partial_tiles(int128):
    # 1, 17 and 33 queries: one partial tile, a tile and a bit, two tiles and one
    x, q, oi, gi = int128
    gi.nprobe = 6
    for n in (1, 17, 33):
        Dg, Ig = tile_search(gi, q[:n], 10, 60)
        Do, Io = oi.search(q[:n], 10, 6, k_factor=60)
        assert_same(Dg, Ig, Do, Io)


@pytest.mark.parametrize("d,M,nlist,nprobe", [(96, 12, 160, 100), (96, 12, 64, 20), (64, 8, 50, 10)])
def test_tile_scan_other_shapes(gpu, d, M, nlist, nprobe):
    # M = 12 (d = 96, the 10M x 96 layout) with more than 64 probes (16-query
    # tiles) and with fewer (32-query tiles); M = 8 (d = 64)
    x = clustered(720 + d, 15000, d, nclusters=50)
    x[0, :], x[1, :] = 0.0, 255.0
    oi, gi = int_index(gpu, x, d, nlist, M, 721 + d)
    oi.add(x)
    gi.add(x)
    q = clustered(722 + d, 70, d, nclusters=50)
    gi.nprobe = nprobe
    for k, kf in ((10, 100), (100, 3)):
        Dg, Ig = tile_search(gi, q, k, kf)
        Do, Io = oi.search(q, k, nprobe, k_factor=kf)
        assert_same(Dg, Ig, Do, Io)


def test_tile_scan_ties_break_by_id(gpu):
    # every row stored three times: each distance appears (at least) three
    # times, so the k * k_factor cut falls inside runs of equal distances
    base = clustered(730, 3000, 128, nclusters=30)
    x = np.repeat(base, 3, axis=0)
    x[0, :], x[1, :] = 0.0, 255.0
    oi, gi = int_index(gpu, x, 128, 20, 16, 731)
    oi.add(x)
    gi.add(x)
    q = clustered(732, 40, 128, nclusters=30)
    gi.nprobe = 5
    for k, kf in ((10, 31), (7, 100), (50, 11)):
        Dg, Ig = tile_search(gi, q, k, kf)
        Do, Io = oi.search(q, k, 5, k_factor=kf)
        assert_same(Dg, Ig, Do, Io)


@pytest.mark.parametrize("copies", [3000, 70000])
def test_tile_scan_crowded_bin_falls_back(gpu, copies):
    # one row repeated `copies` times: the kk-th distance's bin holds more
    # keys than the boundary buffer (3000) or than a 16-bit bin counts
    # (70000); those queries go to the FP32-table deep-pool scan, exact here
    rng = np.random.default_rng(740)
    other = clustered(741, 3000, 128, nclusters=20)
    dup = np.repeat(other[:1], copies, axis=0)
    x = np.concatenate([other, dup])
    x[0, :], x[1, :] = 0.0, 255.0
    oi, gi = int_index(gpu, x, 128, 8, 16, 742)
    oi.add(x)
    gi.add(x)
    q = np.concatenate([other[:1] + rng.integers(-2, 3, (8, 128)), clustered(743, 8, 128, nclusters=20)])
    q = np.clip(q, 0, 255).astype(np.float32)
    gi.nprobe = 3
    Dg, Ig = tile_search(gi, q, 10, 300)
    Do, Io = oi.search(q, 10, 3, k_factor=300)
    assert_same(Dg, Ig, Do, Io)


# --------------------------------------------------------- float data
@pytest.fixture(scope="module")
def float128(gpu):
    rng = np.random.default_rng(750)
    x = (clustered(751, 24000, 128, nclusters=120) + rng.uniform(-0.5, 0.5, (24000, 128))).astype(np.float32)
    q = (clustered(752, 150, 128, nclusters=120) + rng.uniform(-0.5, 0.5, (150, 128))).astype(np.float32)
    oi = o.IVFPQIndex(128, 48, 16, 0)
    oi.train(x, niter_coarse=3, niter_pq=3)
    coarse, pq = oi.codebooks()
    oi.add(x)
    gi = gpu.IndexIVFPQ(128, 48, 16, 0)
    gi.set_codebooks(coarse, pq)
    gi.add(x)
    return x, q, oi, gi


def test_tile_scan_every_probed_entry_is_exact(float128):
    # every probed entry a candidate: the FP32 rerank returns the exact
    # search over the probed rows, which the oracle computes
    x, q, oi, gi = float128
    sizes = np.sort(gi.list_sizes())[::-1]
    nprobe = int(np.searchsorted(np.cumsum(sizes), 4090, side="right"))
    kf = int(math.ceil(sizes[:nprobe].sum() / 10))
    assert nprobe >= 2 and 10 * kf <= 4096
    gi.nprobe = nprobe
    gi.set_prefilter(0)
    try:
        Dg, Ig = tile_search(gi, q, 10, kf)
    finally:
        gi.set_prefilter(8)
    Do, Io = oi.search(q, 10, nprobe, k_factor=kf)
    assert_same(Dg, Ig, Do, Io)


def test_tile_scan_recall_matches_fp32_tables(float128):
    # the measured pool width (k = 10, k_factor 288) on float rows: the FP16
    # tile distances choose candidates as well as the FP32 tables (recall of
    # each against the other >= 0.99; a row both return has the same exact
    # rerank distance) and within 0.005 of them against the FP64 brute force
    x, q, oi, gi = float128
    gi.nprobe = 16
    gi.set_search_mode("mixed")
    try:
        D32, I32 = gi.search(q, 10, k_factor=288)
    finally:
        gi.set_search_mode("plain")
    D16, I16 = tile_search(gi, q, 10, 288)
    assert recall(I32, I16) >= 0.99
    for i in range(q.shape[0]):
        d32 = dict(zip(I32[i].tolist(), D32[i].tolist()))
        for j, dv in zip(I16[i].tolist(), D16[i].tolist()):
            if j in d32:
                assert d32[j] == dv
    x64, q64 = x.astype(np.float64), q.astype(np.float64)
    d2 = (q64 ** 2).sum(1)[:, None] + (x64 ** 2).sum(1)[None] - 2 * q64 @ x64.T
    gt = np.argsort(d2, axis=1, kind="stable")[:, :10]
    assert recall(gt, I16) >= recall(gt, I32) - 0.005


def test_tile_scan_batch_composition_does_not_change_results(float128):
    # a query's candidates depend on its own distances only: searched in a
    # batch of 150, alone, or in a shuffled batch, it returns the same result
    x, q, oi, gi = float128
    gi.nprobe = 16
    Db, Ib = tile_search(gi, q, 10, 288)
    perm = np.random.default_rng(760).permutation(q.shape[0])
    Dp, Ip = tile_search(gi, q[perm], 10, 288)
    assert_same(Dp, Ip, Db[perm], Ib[perm])
    for i in (0, 77, 149):
        Da, Ia = tile_search(gi, q[i:i + 1], 10, 288)
        assert_same(Da, Ia, Db[i:i + 1], Ib[i:i + 1])


def test_tile_scan_graph_replays_equal_eager(float128):
    import torch

    x, q, oi, gi = float128
    gi.nprobe = 16
    gi.set_search_mode("mixed")
    gi.set_table_bits(16)
    gi.set_tile_min_batch(1)
    try:
        De, Ie = gi.search(q, 10, k_factor=288)
        dev = torch.device("cuda:0")
        qd = torch.from_numpy(q).to(dev)
        D = torch.empty((q.shape[0], 10), dtype=torch.float32, device=dev)
        I = torch.empty((q.shape[0], 10), dtype=torch.int64, device=dev)
        g = gi.capture(qd.data_ptr(), q.shape[0], 10, D.data_ptr(), I.data_ptr(), k_factor=288)
        for rep in range(3):
            D.fill_(0.0)
            I.fill_(0)
            torch.cuda.synchronize()
            g.launch()
            torch.cuda.synchronize()
            assert_same(D.cpu().numpy(), I.cpu().numpy(), De, Ie)
            gi.search(q[:5], 10, k_factor=288)  # eager work between replays
        del g
    finally:
        gi.set_tile_min_batch(0)
        gi.set_table_bits(32)
        gi.set_search_mode("plain")


def test_tile_scan_small_batches_take_the_8bit_scan(float128):
    # below the default minimum (16 queries per CU) table bits 16 runs the
    # 8-bit deep-pool scan: the same results as table bits 8
    x, q, oi, gi = float128
    gi.nprobe = 16
    gi.set_search_mode("mixed")
    try:
        gi.set_table_bits(8)
        D8, I8 = gi.search(q, 10, k_factor=288)
        gi.set_table_bits(16)
        D16, I16 = gi.search(q, 10, k_factor=288)
    finally:
        gi.set_table_bits(32)
        gi.set_search_mode("plain")
    assert_same(D16, I16, D8, I8)


@pytest.mark.parametrize("kf", [6, 40])
def test_tile_scan_per_query_k_batch_equals_per_k_searches(int128, kf):
    # k in {1, 10, 100} and nprobe in 4..24 per query in one batch on tiles
    # (scanned for the largest k * k_factor, each query then cut to its own
    # k * k_factor) equals each k group searched alone; the captured graph
    # of the batch replays the same results
    import torch

    x, q, oi, gi = int128
    rng = np.random.default_rng(770 + kf)
    kq = rng.choice([1, 10, 100], q.shape[0]).astype(np.int32)
    npq = rng.integers(4, 25, q.shape[0]).astype(np.int32)
    kmax = int(kq.max())
    gi.set_search_mode("mixed")
    gi.set_table_bits(16)
    gi.set_tile_min_batch(1)
    try:
        De = np.full((q.shape[0], kmax), np.float32(1e38))
        Ie = np.full((q.shape[0], kmax), -1, np.int64)
        for kv in np.unique(kq):
            sel = kq == kv
            Dk, Ik = gi.search(q[sel], int(kv), k_factor=kf, nprobe_per_query=npq[sel])
            De[sel, :kv], Ie[sel, :kv] = Dk, Ik
        dev = torch.device("cuda:0")
        qd = torch.from_numpy(q).to(dev)
        kd = torch.from_numpy(kq).to(dev)
        nd = torch.from_numpy(npq).to(dev)
        D = torch.empty((q.shape[0], kmax), dtype=torch.float32, device=dev)
        I = torch.empty((q.shape[0], kmax), dtype=torch.int64, device=dev)
        gi.search_device(qd.data_ptr(), q.shape[0], kmax, D.data_ptr(), I.data_ptr(), k_factor=kf,
                         nprobe_ptr=nd.data_ptr(), max_nprobe=24, kq_ptr=kd.data_ptr())
        torch.cuda.synchronize()
        assert_same(D.cpu().numpy(), I.cpu().numpy(), De, Ie)
        g = gi.capture(qd.data_ptr(), q.shape[0], kmax, D.data_ptr(), I.data_ptr(), k_factor=kf,
                       nprobe_ptr=nd.data_ptr(), max_nprobe=24, kq_ptr=kd.data_ptr())
        for rep in range(2):
            D.fill_(0.0)
            I.fill_(0)
            torch.cuda.synchronize()
            g.launch()
            torch.cuda.synchronize()
            assert_same(D.cpu().numpy(), I.cpu().numpy(), De, Ie)
        del g
    finally:
        gi.set_tile_min_batch(0)
        gi.set_table_bits(32)
        gi.set_search_mode("plain")
