"""Mixed-precision Flat candidates (BF16 and FP8 E4M3 on MFMA, FP32
accumulate) with the exact FP32 rerank, against the oracle and FP64.

The FP8 row encoding is checked byte for byte against an independent numpy
encoder of OCP E4M3 (bias 7, max 448, subnormals, round to nearest even) under
the documented per-row power-of-two scaling. With a rerank over every row the
mixed forms must return the exact search; with a shallow rerank they must keep
recall and state their distance error against FP64.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as o

pytestmark = pytest.mark.gpu


def e4m3_table():
    """Values of the 256 OCP E4M3 codes (NaN for S.1111.111)."""
    vals = np.empty(256, np.float64)
    for c in range(256):
        s = -1.0 if c & 0x80 else 1.0
        e = (c >> 3) & 0xF
        m = c & 7
        if e == 15 and m == 7:
            vals[c] = np.nan
        elif e == 0:
            vals[c] = s * m / 8.0 * 2.0 ** -6
        else:
            vals[c] = s * (1 + m / 8.0) * 2.0 ** (e - 7)
    return vals


def e4m3_encode(x):
    """Nearest E4M3 code, ties to the even code (|x| <= 448)."""
    vals = e4m3_table()
    pos = np.arange(0x7F)  # finite non-negative codes
    pv = vals[pos]
    flat = x.reshape(-1).astype(np.float64)
    out = np.empty(flat.size, np.uint8)
    for i, v in enumerate(flat):
        dist = np.abs(pv - abs(v))
        best = pos[dist == dist.min()]
        if best.size > 1:  # a tie between neighbours: the even code
            best = best[(best & 1) == 0]
        c = int(best[0])
        out[i] = c | 0x80 if (v < 0 and c != 0) else c
    return out.reshape(x.shape)


def row_scale_exponent(row):
    mx = np.abs(row).max()
    if mx == 0:
        return 0
    ex = int(np.frexp(np.float32(mx))[1])
    e = 9 - ex
    if np.ldexp(np.float32(mx), e) > 448:
        e -= 1
    return e


def test_fp8_encoding_matches_ocp_e4m3(gpu):
    from mojovec_amd import _lib

    rng = np.random.default_rng(3)
    n, d = 40, 72
    x = (rng.standard_normal((n, d)) * rng.choice([1e-3, 0.05, 1.0, 30.0], (n, 1))).astype(np.float32)
    x[0] = 0.0
    x[1, :5] = [448 / 512, -0.1, 1e-9, 3e-3, 0.5]  # ties, subnormals, underflow
    rb = (d + 63) // 64 * 64
    codes = np.zeros((n, rb), np.uint8)
    inv = np.zeros(n, np.float32)
    norms = np.zeros(n, np.float32)
    L = _lib.load()
    _lib.check(L.mvg_fp8_encode(x.ctypes.data_as(ctypes.c_void_p), n, d, codes.ctypes.data_as(ctypes.c_void_p),
                                inv.ctypes.data_as(ctypes.c_void_p), norms.ctypes.data_as(ctypes.c_void_p)))
    vals = e4m3_table()
    for r in range(n):
        e = row_scale_exponent(x[r])
        assert inv[r] == np.float32(2.0 ** -e)
        scaled = np.ldexp(x[r].astype(np.float64), e)
        assert np.abs(scaled).max() <= 448
        exp_codes = e4m3_encode(scaled)
        got = codes[r, :d]
        # +0 and -0 both decode to zero; compare values, then the bytes
        assert np.array_equal(vals[got], vals[exp_codes]), f"row {r}"
        assert (codes[r, d:] == 0).all()
        deq = vals[got] * 2.0 ** -e
        assert abs(float(norms[r]) - float((deq ** 2).sum())) <= 1e-5 * max(1.0, float((deq ** 2).sum()))


@pytest.mark.parametrize("form", ["mixed-bf16", "mixed-fp8"])
@pytest.mark.parametrize("metric", [0, 1])
def test_mixed_flat_with_full_rerank_is_exact(gpu, form, metric):
    # integer data: every FP32 order is exact, so reranking every row gives
    # the plain search bit for bit
    rng = np.random.default_rng(11 + metric)
    x = rng.integers(0, 64, (500, 128)).astype(np.float32)
    q = rng.integers(0, 64, (30, 128)).astype(np.float32)
    ix = gpu.IndexFlat(128, metric)
    ix.add(x)
    Dp, Ip = ix.search(q, 2)
    ix.set_search_mode(form, k_factor=250)  # k * k_factor = every row
    Dm, Im = ix.search(q, 2)
    assert np.array_equal(Ip, Im) and np.array_equal(Dp, Dm)
    Do, Io = o.flat_search(x, q, 2, metric)
    assert np.array_equal(Im, Io)


@pytest.mark.parametrize("form,bar", [("mixed-bf16", 0.97), ("mixed-fp8", 0.9)])
@pytest.mark.parametrize("metric", [0, 1])
def test_mixed_flat_recall_and_error_vs_fp64(gpu, form, bar, metric):
    # unit-norm embeddings (the C3 recipe at small scale)
    rng = np.random.default_rng(21 + metric)
    d, n, nq = 256, 20000, 200
    dirs = rng.standard_normal((200, d))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)

    def draw(m):
        v = dirs[rng.integers(0, 200, m)] + 0.3 / np.sqrt(d) * rng.standard_normal((m, d))
        return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)

    x, q = draw(n), draw(nq)
    x64, q64 = x.astype(np.float64), q.astype(np.float64)
    dd = (-(q64 @ x64.T)) if metric == 1 else ((q64 ** 2).sum(1)[:, None] + (x64 ** 2).sum(1)[None] - 2 * q64 @ x64.T)
    gt = np.argsort(dd, axis=1, kind="stable")[:, :10]
    ix = gpu.IndexFlat(d, metric)
    ix.add(x)
    ix.set_search_mode(form, k_factor=4)
    D, I = ix.search(q, 10)
    recall = np.mean([len(set(a) & set(b)) / 10 for a, b in zip(gt.tolist(), I.tolist())])
    assert recall >= bar, recall
    # the reranked distances are exact FP32 of the returned rows
    exact = (q64 * x64[I[:, 0]]).sum(1) if metric == 1 else ((q64 - x64[I[:, 0]]) ** 2).sum(1)
    assert np.max(np.abs(D[:, 0] - exact)) < 1e-4


@pytest.mark.parametrize("form", ["mixed-bf16", "mixed-fp8"])
@pytest.mark.parametrize("metric", [0, 1])
def test_mixed_candidates_exact_on_small_integers(gpu, form, metric):
    # Integers in [-15, 15] are exact in bf16 and in FP8 E4M3 under the
    # power-of-two row scaling (4 significant bits), and every product sum is
    # exact in FP32, so the MFMA candidate distances equal the plain FP32
    # search, ties (lower id first) and exclusions included. 300 queries span
    # several 64-query workgroup tiles, the last one partial; k = 64 is the
    # widest candidate set the 64-query form keeps.
    rng = np.random.default_rng(50 + metric)
    lo = -15 if metric == 1 else 0
    x = rng.integers(lo, 16, (20000, 64)).astype(np.float32)
    q = rng.integers(lo, 16, (300, 64)).astype(np.float32)
    ex = (rng.random(20000) < 0.2).astype(np.uint8)
    ix = gpu.IndexFlat(64, metric)
    ix.add(x)
    for k in (1, 10, 64):
        ix.set_search_mode("plain")
        Dp, Ip = ix.search(q, k, exclude=ex)
        ix.set_search_mode(form)
        Dm, Im = ix.search(q, k, exclude=ex)
        assert np.array_equal(Im, Ip), f"k={k}"
        assert np.array_equal(Dm, Dp), f"k={k}"
    Do, Io = o.flat_search(x, q, 10, metric, exclude=ex)
    ix.set_search_mode(form)
    assert np.array_equal(ix.search(q, 10, exclude=ex)[1], Io)
