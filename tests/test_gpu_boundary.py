"""GPU checks of the boundary's later entry points against the oracle:
IndexIVFPQ.add_with_ids, the Flat / SQ8 row readback, ProductQuantizer.decode
on the device, the per-query-k batch search and its captured graph, and the
state guards (a graph refuses to replay after its index changed; a served
index refuses mutations and setting changes)."""
import numpy as np
import pytest

from oracle import oracle as o

pytestmark = pytest.mark.gpu


def clustered(seed, n, d, nclusters=40, sigma=12.0):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0, 128, (nclusters, d))
    lab = rng.integers(0, nclusters, n)
    x = centres[lab] + rng.normal(0, sigma, (n, d))
    return np.rint(np.clip(x, 0, 255)).astype(np.float32)


def assert_same(Dg, Ig, Do, Io):
    assert np.array_equal(Ig, Io), f"id mismatch at {np.argwhere(Ig != Io)[:5].tolist()}"
    assert np.array_equal(Dg, Do), f"distance mismatch at {np.argwhere(Dg != Do)[:5].tolist()}"


def _pair(gpu, x, d, nlist, M, metric, niter=3):
    oi = o.IVFPQIndex(d, nlist, M, metric)
    oi.train(x, niter_coarse=niter, niter_pq=niter)
    coarse, pq = oi.codebooks()
    gi = gpu.IndexIVFPQ(d, nlist, M, metric)
    gi.set_codebooks(coarse, pq)
    return oi, gi


# ------------------------------------------------------------ add_with_ids
@pytest.mark.parametrize("metric", [0, 1])
def test_ivfpq_add_with_ids_matches_oracle(gpu, metric):
    # unordered ids up to 10^6, in two adds, then one sequential add; plain
    # and mixed forms, with the rerank (rows found through the id map), the
    # SQ8 pre-rerank and exclusions by id
    x = clustered(501 + metric, 7000, 32)
    if metric == 1:
        x = x / 255.0
    rng = np.random.default_rng(502)
    ids = rng.permutation(1000000)[:6000].astype(np.int64) + 7000  # clear of the sequential ids
    oi, gi = _pair(gpu, x, 32, 24, 8, metric)
    for ix in (oi, gi):
        ix.add_with_ids(x[:2500], ids[:2500])
        ix.add_with_ids(x[2500:6000], ids[2500:])
    # a sequential add continues from ntotal (6000): ids 6000..6999
    oi.add(x[6000:])
    gi.add(x[6000:])
    assert gi.ntotal == 7000
    for l in range(24):
        io, co = oi.list(l)
        ig, cg = gi.get_list(l)
        assert np.array_equal(ig, io) and np.array_equal(cg, co)
    q = clustered(503, 60, 32)
    if metric == 1:
        q = q / 255.0
    all_ids = np.concatenate([ids, np.arange(6000, 7000)])
    ex = np.zeros(int(all_ids.max()) + 1, np.uint8)
    ex[all_ids[::9]] = 1
    for nprobe, k, kf in [(4, 10, 1), (6, 10, 4), (24, 5, 120), (8, 700, 1)]:
        gi.nprobe = nprobe
        Dg, Ig = gi.search(q, k, k_factor=kf, exclude=ex)
        Do, Io = oi.search(q, k, nprobe, k_factor=kf, exclude=ex)
        assert_same(Dg, Ig, Do, Io)
    if metric == 0:
        sizes = np.sort(gi.list_sizes())[::-1]
        nprobe = int(np.searchsorted(np.cumsum(sizes), 4000, side="right"))
        kf = int(np.ceil(sizes[:nprobe].sum() / 4))
        gi.set_search_mode("mixed")
        try:
            gi.nprobe = nprobe
            gi.set_prefilter(0)
            Dm, Im = gi.search(q, 4, k_factor=kf)  # every probed entry reranked: exact
            Do, Io = oi.search(q, 4, nprobe, k_factor=kf)
            assert_same(Dm, Im, Do, Io)
            gi.set_prefilter(8)
            Dp, Ip = gi.search(q, 4, k_factor=kf)
            same = Ip == Io
            assert same.mean() >= 0.99 and np.array_equal(Dp[same], Do[same])
        finally:
            gi.set_search_mode("plain")


def test_ivfpq_add_with_ids_rejects_bad_and_duplicate_ids(gpu):
    x = clustered(510, 600, 16)
    _, gi = _pair(gpu, x, 16, 4, 4, 0)
    with pytest.raises(gpu.MojoVecError, match=r"\[0, 2\^31 - 1\]"):
        gi.add_with_ids(x[:2], np.array([5, -1]))
    gi.add_with_ids(x[:100], np.arange(100) * 3)
    with pytest.raises(gpu.MojoVecError, match="duplicate IVF-PQ id 3"):
        gi.add_with_ids(x[100:102], np.array([1000, 3]))
    assert gi.ntotal == 100  # refused before anything changed
    with pytest.raises(gpu.MojoVecError, match="duplicate"):
        gi.add_with_ids(x[100:102], np.array([7, 7]))
    # without kept rows the reference's behaviour: duplicates are stored
    gn = gpu.IndexIVFPQ(16, 4, 4, keep_vectors=False)
    gn.set_codebooks(*gi.codebooks())
    gn.add_with_ids(x[:3], np.array([9, 9, 9]))
    assert gn.ntotal == 3
    gn.nprobe = 4
    _, I = gn.search(x[:1], 3)
    assert (I == 9).all()


# ------------------------------------------------------------ row readback
def test_flat_and_sq8_get_vector(gpu):
    rng = np.random.default_rng(520)
    x = rng.standard_normal((3000, 40)).astype(np.float32)
    f = gpu.IndexFlat(40)
    f.add(x[:1000])
    f.add(x[1000:])
    s = gpu.IndexFlatSQ8(40, 1)
    s.add(x)
    ids = rng.integers(0, 3000, 257)
    for ix in (f, s):
        assert np.array_equal(ix.get_vectors(ids), x[ids])
        assert np.array_equal(ix.get_vector(2999), x[2999])
        assert ix.get_vectors(np.zeros(0, np.int64)).shape == (0, 40)
        with pytest.raises(gpu.MojoVecError, match=r"\[0, ntotal\)"):
            ix.get_vector(3000)
        with pytest.raises(gpu.MojoVecError, match=r"\[0, ntotal\)"):
            ix.get_vectors(np.array([0, -1]))


def test_pq_decode_matches_oracle(gpu):
    rng = np.random.default_rng(530)
    x = rng.standard_normal((3000, 48)).astype(np.float32)
    for M in (6, 12, 48):
        pq = gpu.ProductQuantizer(48, M, niter=3)
        pq.train(x)
        codes = pq.compute_codes(x[:500])
        got = pq.decode(codes)
        assert np.array_equal(got, o.pq_decode(codes, pq.centroids))
        # the ADC identity of tests/test_pq.mojo:57-114 through the decode
        lut = pq.compute_distance_table(x[:3])
        adc = np.array([[lut[i, m, codes[j, m]] for m in range(M)] for i in range(3) for j in range(5)]).sum(1)
        rec = ((x[:3, None].astype(np.float64) - got[None, :5]) ** 2).sum(-1).reshape(-1)
        assert np.allclose(adc, rec, atol=1e-3, rtol=1e-5)


# -------------------------------------------------------- per-query k
@pytest.fixture(scope="module")
def kq_index(gpu):
    x = clustered(540, 20000, 64, nclusters=60)
    x[0, :], x[1, :] = 0.0, 255.0
    _, gi = _pair(gpu, x, 64, 48, 16, 0)
    gi.add(x)
    q = clustered(541, 300, 64, nclusters=60)
    rng = np.random.default_rng(542)
    kq = rng.choice([1, 10, 100], q.shape[0]).astype(np.int32)
    npq = rng.integers(4, 49, q.shape[0]).astype(np.int32)
    return gi, q, kq, npq


def _per_k_eager(gi, q, kq, npq, kf, kmax):
    D = np.full((q.shape[0], kmax), np.float32(1e38))
    I = np.full((q.shape[0], kmax), -1, np.int64)
    for kv in np.unique(kq):
        sel = kq == kv
        Dk, Ik = gi.search(q[sel], int(kv), k_factor=kf, nprobe_per_query=npq[sel])
        D[sel, :kv], I[sel, :kv] = Dk, Ik
    return D, I


@pytest.mark.parametrize("mode,kf", [(m, kf) for m in ("plain", "plain-wq", "plain-split", "mixed", "mixed-deep")
                                      for kf in (1, 6)] + [("mixed", 40), ("mixed-deep", 40), ("plain", 40)])
def test_per_query_k_batch_equals_per_k_searches(gpu, kq_index, mode, kf):
    # k in {1, 10, 100} per query, nprobe 4..48 per query: one batch scanned
    # for the largest k * k_factor equals each k group searched alone. With
    # k_factor 40 the mixed batch takes the SQ8 pre-rerank at each query's
    # own width (80 of 400 for k = 10, none for k = 1).
    import torch

    gi, q, kq, npq = kq_index
    kmax = int(kq.max())
    gi.set_search_mode("mixed" if mode.startswith("mixed") else "plain", split_scan=mode == "plain-split")
    gi.set_work_queue(1024 if mode == "plain-wq" else 0)
    gi.set_deep_select(1 if mode == "mixed-deep" else 0)
    try:
        De, Ie = _per_k_eager(gi, q, kq, npq, kf, kmax)
        dev = torch.device("cuda:0")
        qd = torch.from_numpy(q).to(dev)
        kd = torch.from_numpy(kq).to(dev)
        nd = torch.from_numpy(npq).to(dev)
        D = torch.empty((q.shape[0], kmax), dtype=torch.float32, device=dev)
        I = torch.empty((q.shape[0], kmax), dtype=torch.int64, device=dev)
        gi.search_device(qd.data_ptr(), q.shape[0], kmax, D.data_ptr(), I.data_ptr(), k_factor=kf,
                         nprobe_ptr=nd.data_ptr(), max_nprobe=48, kq_ptr=kd.data_ptr())
        torch.cuda.synchronize()
        assert_same(D.cpu().numpy(), I.cpu().numpy(), De, Ie)
    finally:
        gi.set_search_mode("plain")
        gi.set_work_queue(0)
        gi.set_deep_select(0)


def test_per_query_k_graph_replays_and_refuses_after_a_change(gpu, kq_index):
    import torch

    gi, q, kq, npq = kq_index
    kmax, kf = int(kq.max()), 2
    dev = torch.device("cuda:0")
    qd = torch.from_numpy(q).to(dev)
    kd = torch.from_numpy(kq).to(dev)
    nd = torch.from_numpy(npq).to(dev)
    D = torch.empty((q.shape[0], kmax), dtype=torch.float32, device=dev)
    I = torch.empty((q.shape[0], kmax), dtype=torch.int64, device=dev)
    g = gi.capture(qd.data_ptr(), q.shape[0], kmax, D.data_ptr(), I.data_ptr(), k_factor=kf,
                   nprobe_ptr=nd.data_ptr(), max_nprobe=48, kq_ptr=kd.data_ptr())
    for rep in range(3):
        qn = clustered(550 + rep, q.shape[0], 64, nclusters=60)
        qd.copy_(torch.from_numpy(qn))
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        De, Ie = _per_k_eager(gi, qn, kq, npq, kf, kmax)
        assert_same(D.cpu().numpy(), I.cpu().numpy(), De, Ie)
    # a mutation moves the index on: the graph no longer replays
    gi2 = gpu.IndexIVFPQ(64, 48, 16)
    gi2.set_codebooks(*gi.codebooks())
    x = clustered(540, 2000, 64, nclusters=60)
    gi2.add(x)
    g2 = gi2.capture(qd.data_ptr(), 20, 10, D.data_ptr(), I.data_ptr())
    g2.launch()
    torch.cuda.synchronize()
    gi2.add(x[:10])
    with pytest.raises(gpu.MojoVecError, match="changed after this graph was captured"):
        g2.launch()
    del g2


def test_served_index_refuses_changes(gpu):
    x = clustered(560, 3000, 32)
    _, gi = _pair(gpu, x, 32, 16, 8, 0)
    gi.add(x)
    srv = gpu.Server(gi, slots=2, max_batch=256, mem_budget=64 << 20)
    try:
        for call in (lambda: gi.set_search_mode("mixed"), lambda: gi.set_prefilter(0),
                     lambda: gi.set_work_queue(512), lambda: gi.set_deep_select(1),
                     lambda: gi.add(x[:5])):
            with pytest.raises(gpu.MojoVecError, match="held by a server"):
                call()
        t = srv.submit(x[:7], k=5, nprobe=4)
        D, I, _ = srv.wait(t)
        gi.nprobe = 4
        Da, Ia = gi.search(x[:7], 5)  # eager searches still run beside the server
        assert_same(D, I, Da, Ia)
    finally:
        srv.close()
    gi.set_search_mode("plain")  # released with the server
    gi.add(x[:5])
    assert gi.ntotal == 3005
