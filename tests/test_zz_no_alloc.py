"""The allocation-free IVF-PQ device search (mvg_ivfpq_search_device_noalloc)
through the C ABI: on two calls in a row, the second with a smaller batch,
its results equal mvg_ivfpq_search_device's bit for bit in the plain and
mixed forms, the work-queue and deep-pool schedules and with per-query
nprobe; a request beyond the 64 MiB scratch and an unprepared index are
refused with their messages."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def clustered(seed, n, d, nclusters=40, sigma=12.0):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0, 128, (nclusters, d))
    lab = rng.integers(0, nclusters, n)
    x = centres[lab] + rng.normal(0, sigma, (n, d))
    return np.rint(np.clip(x, 0, 255)).astype(np.float32)


def _vp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _run(lib, fn, ix, q, k, kf, npq, maxp, ex=None):
    import torch

    nq = q.shape[0]
    D = torch.full((nq, k), -7.0, dtype=torch.float32, device="cuda:0")
    I = torch.full((nq, k), -7, dtype=torch.int64, device="cuda:0")
    rc = fn(ix._h, _vp(q), nq, k, _vp(npq), maxp, kf, _vp(ex), 0 if ex is None else ex.numel(),
            _vp(D), _vp(I), None)
    assert rc == 0, lib.mvg_last_error().decode()
    torch.cuda.synchronize()
    return D.cpu().numpy(), I.cpu().numpy()


@pytest.fixture(scope="module")
def index(gpu):
    x = clustered(700, 20000, 64, nclusters=50)
    ix = gpu.IndexIVFPQ(64, 64, 16)
    ix.train(x, niter_coarse=3, niter_pq=3)
    ix.add(x)
    return ix


@pytest.mark.parametrize("mode", ["plain", "mixed", "plain-wq", "plain-deep", "mixed-deep"])
def test_noalloc_equals_search_device_twice(gpu, index, mode):
    import torch

    from mojovec_amd import _lib

    lib = _lib.load()
    ix = index
    ix.set_search_mode("mixed" if mode.startswith("mixed") else "plain")
    ix.set_work_queue(1024 if mode == "plain-wq" else 0)
    ix.set_deep_select(1 if mode.endswith("deep") else 0)
    try:
        ix.search(np.zeros((1, 64), np.float32), 1)  # prepares the index for this form
        rng = np.random.default_rng(701)
        for nq, k, kf in ((400, 10, 8), (37, 5, 2)):  # second call smaller
            q = torch.from_numpy(clustered(702 + nq, nq, 64, nclusters=50)).to("cuda:0")
            npq = torch.from_numpy(rng.integers(1, 40, nq).astype(np.int32)).to("cuda:0")
            ex = torch.from_numpy((rng.random(20000) < 0.1).astype(np.uint8)).to("cuda:0")
            Dr, Ir = _run(lib, lib.mvg_ivfpq_search_device, ix, q, k, kf, npq, 40, ex)
            Dn, In = _run(lib, lib.mvg_ivfpq_search_device_noalloc, ix, q, k, kf, npq, 40, ex)
            assert np.array_equal(In, Ir) and np.array_equal(Dn, Dr), (mode, nq)
            assert (In[:, 0] >= 0).all()
    finally:
        ix.set_search_mode("plain")
        ix.set_work_queue(0)
        ix.set_deep_select(0)


def test_noalloc_refuses_oversized_and_unprepared(gpu, index):
    import torch

    from mojovec_amd import _lib

    lib = _lib.load()
    ix = index
    ix.search(np.zeros((1, 64), np.float32), 1)
    nq = 300000  # probe lists alone: 300000 x 64 x 4 bytes > 64 MiB
    q = torch.zeros((nq, 64), dtype=torch.float32, device="cuda:0")
    D = torch.empty((nq, 10), dtype=torch.float32, device="cuda:0")
    I = torch.empty((nq, 10), dtype=torch.int64, device="cuda:0")
    ix.nprobe = 64
    rc = lib.mvg_ivfpq_search_device_noalloc(ix._h, _vp(q), nq, 10, None, 0, 1, None, 0, _vp(D), _vp(I),
                                             None)
    assert rc == _lib.MVG_ERR_INVALID
    assert b"64 MiB scratch" in lib.mvg_last_error()
    fresh = gpu.IndexIVFPQ(64, 64, 16)  # lists added but never uploaded
    fresh.set_codebooks(*ix.codebooks())
    fresh.add(clustered(710, 500, 64, nclusters=50))
    rc = lib.mvg_ivfpq_search_device_noalloc(fresh._h, _vp(q), 10, 10, None, 0, 1, None, 0, _vp(D),
                                             _vp(I), None)
    assert rc == _lib.MVG_ERR_STATE
    assert b"not prepared" in lib.mvg_last_error()
    fresh.search(np.zeros((1, 64), np.float32), 1)  # prepared now
    rc = lib.mvg_ivfpq_search_device_noalloc(fresh._h, _vp(q), 10, 10, None, 0, 1, None, 0, _vp(D),
                                             _vp(I), None)
    assert rc == 0, lib.mvg_last_error().decode()
    torch.cuda.synchronize()
