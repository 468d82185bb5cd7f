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
