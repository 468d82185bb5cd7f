"""Continuous-batching server: every request's result equals the same request
searched alone, whatever batch it shared and however batches overlapped."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_server_results_match_alone(gpu):
    rng = np.random.default_rng(5)
    x = np.rint(np.clip(rng.normal(100, 30, (20000, 32)), 0, 255)).astype(np.float32)
    ix = gpu.IndexIVFPQ(32, 64, 8)
    ix.train(x, niter_coarse=5, niter_pq=5)
    ix.add(x)
    srv = gpu.Server(ix, slots=3, max_batch=700, mem_budget=64 << 20)
    reqs = []
    for i in range(40):
        nq = int(rng.integers(1, 300))
        q = np.rint(np.clip(rng.normal(100, 30, (nq, 32)), 0, 255)).astype(np.float32)
        k = int(rng.choice([1, 10, 50]))
        nprobe = int(rng.integers(1, 64))
        kf = int(rng.choice([1, 2]))
        reqs.append((q, k, nprobe, kf, srv.submit(q, k=k, nprobe=nprobe, k_factor=kf)))
    outs = [(q, k, nprobe, kf, srv.wait(t)) for q, k, nprobe, kf, t in reqs]
    st = srv.stats()
    assert st["completed_queries"] == sum(q.shape[0] for q, *_ in reqs)
    for q, k, nprobe, kf, (D, I, lat) in outs:
        ix.nprobe = nprobe
        Da, Ia = ix.search(q, k, k_factor=kf)
        assert np.array_equal(I, Ia) and np.array_equal(D, Da)
        assert lat >= 0.0
    srv.close()


@pytest.mark.parametrize("wq", [0, 1024, "deep", "deep-hbm"])
def test_captured_search_graph_replays_match_eager(gpu, wq):
    import torch

    rng = np.random.default_rng(7)
    x = np.rint(np.clip(rng.normal(100, 30, (20000, 32)), 0, 255)).astype(np.float32)
    ix = gpu.IndexIVFPQ(32, 64, 8)
    ix.train(x, niter_coarse=4, niter_pq=4)
    ix.add(x)
    if wq in ("deep", "deep-hbm"):  # every k * k_factor takes the deep-pool scan
        ix.set_deep_select(1, "hbm" if wq == "deep-hbm" else "lds")
    else:
        ix.set_work_queue(wq)
    nq, k = 300, 10
    dev = torch.device("cuda:0")
    q = torch.zeros((nq, 32), dtype=torch.float32, device=dev)
    npq = torch.from_numpy(rng.integers(1, 40, nq).astype(np.int32)).to(dev)
    D = torch.empty((nq, k), dtype=torch.float32, device=dev)
    I = torch.empty((nq, k), dtype=torch.int64, device=dev)
    g = ix.capture(q.data_ptr(), nq, k, D.data_ptr(), I.data_ptr(), k_factor=2,
                   nprobe_ptr=npq.data_ptr(), max_nprobe=40)
    for rep in range(4):
        qn = np.rint(np.clip(rng.normal(100, 30, (nq, 32)), 0, 255)).astype(np.float32)
        q.copy_(torch.from_numpy(qn))
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        De, Ie = ix.search(qn, k, k_factor=2, nprobe_per_query=npq.cpu().numpy())
        assert np.array_equal(I.cpu().numpy(), Ie), f"replay {rep}"
        assert np.array_equal(D.cpu().numpy(), De), f"replay {rep}"


def test_captured_mixed_search_with_prefilter_replays_match_eager(gpu):
    # mixed form with k * k_factor above the pre-rerank width: the captured
    # graph holds the query SQ8 encode (d = 32 rows padded to 64 bytes), the
    # pre-rerank and the FP32 rerank
    import torch

    rng = np.random.default_rng(17)
    x = np.rint(np.clip(rng.normal(100, 30, (20000, 32)), 0, 255)).astype(np.float32)
    ix = gpu.IndexIVFPQ(32, 64, 8)
    ix.train(x, niter_coarse=4, niter_pq=4)
    ix.add(x)
    ix.set_search_mode("mixed")
    nq, k, kf = 200, 10, 20
    dev = torch.device("cuda:0")
    q = torch.zeros((nq, 32), dtype=torch.float32, device=dev)
    npq = torch.from_numpy(rng.integers(1, 40, nq).astype(np.int32)).to(dev)
    D = torch.empty((nq, k), dtype=torch.float32, device=dev)
    I = torch.empty((nq, k), dtype=torch.int64, device=dev)
    g = ix.capture(q.data_ptr(), nq, k, D.data_ptr(), I.data_ptr(), k_factor=kf,
                   nprobe_ptr=npq.data_ptr(), max_nprobe=40)
    for rep in range(4):
        qn = np.rint(np.clip(rng.normal(100, 30, (nq, 32)), 0, 255)).astype(np.float32)
        q.copy_(torch.from_numpy(qn))
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        De, Ie = ix.search(qn, k, k_factor=kf, nprobe_per_query=npq.cpu().numpy())
        assert np.array_equal(I.cpu().numpy(), Ie), f"replay {rep}"
        assert np.array_equal(D.cpu().numpy(), De), f"replay {rep}"
    ix.set_search_mode("plain")


def test_server_mixed_precision_and_work_queue(gpu):
    rng = np.random.default_rng(6)
    x = rng.standard_normal((10000, 64)).astype(np.float32)
    ix = gpu.IndexIVFPQ(64, 32, 16)
    ix.train(x, niter_coarse=4, niter_pq=4)
    ix.add(x)
    ix.set_search_mode("mixed")
    ix.set_work_queue(512)
    srv = gpu.Server(ix, slots=2, max_batch=256, mem_budget=32 << 20)
    q = rng.standard_normal((500, 64)).astype(np.float32)
    tickets = [srv.submit(q[i:i + 50], k=10, nprobe=8, k_factor=4) for i in range(0, 500, 50)]
    res = [srv.wait(t) for t in tickets]
    ix.nprobe = 8
    Da, Ia = ix.search(q, 10, k_factor=4)
    assert np.array_equal(np.concatenate([r[1] for r in res]), Ia)
    srv.close()


@pytest.mark.parametrize("metric", [0, 1])
@pytest.mark.parametrize("use_mfma", [True, False])
def test_sq8_server_results_match_alone(gpu, metric, use_mfma):
    # requests of 1..300 queries share batches of up to 700; the 16- and
    # 64-query MFMA forms and the split count all change with the batch size
    rng = np.random.default_rng(8 + metric)
    x = rng.standard_normal((6000, 96)).astype(np.float32)
    ix = gpu.IndexFlatSQ8(96, metric)
    ix.add(x)
    ix.set_search_mode(use_mfma)
    srv = gpu.Server(ix, slots=3, max_batch=700, mem_budget=256 << 20)
    reqs = []
    for i in range(30):
        nq = int(rng.integers(1, 300))
        q = rng.standard_normal((nq, 96)).astype(np.float32)
        k = int(rng.choice([1, 10, 50]))
        kf = int(rng.choice([1, 4]))
        reqs.append((q, k, kf, srv.submit(q, k=k, nprobe=1, k_factor=kf)))
    outs = [(q, k, kf, srv.wait(t)) for q, k, kf, t in reqs]
    assert srv.stats()["completed_queries"] == sum(q.shape[0] for q, *_ in reqs)
    for q, k, kf, (D, I, lat) in outs:
        Da, Ia = ix.search(q, k, k_factor=kf)
        assert np.array_equal(I, Ia) and np.array_equal(D, Da), (q.shape[0], k, kf)
    srv.close()
