"""Pins the CPU oracle (oracle/mvref.c) before it is trusted as the checker.

Each check compares the oracle with something that does not share its code:
exact integer arithmetic, float64 numpy, the reference tests' known answers
(tests/golden, from MojoVec's tests/*.mojo) or invariants the mathematics
fixes. Where the reference leaves an arithmetic detail to the (unvendored)
Mojo stdlib -- reduce_add lane order, round() ties, the Philox stream -- the
test says so and pins the restatement's documented choice only.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as o

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rint(rng, shape, lo=0, hi=256):
    return rng.integers(lo, hi, size=shape).astype(np.float32)


# ---------------------------------------------------------------- distances
@pytest.mark.parametrize("W", [4, 8, 64])
@pytest.mark.parametrize("d", [1, 3, 4, 7, 8, 13, 31, 32, 33, 96, 127, 128, 257, 768])
def test_l2_ip_exact_on_integer_data(W, d):
    # Every order of FP32 operations is exact here (d * 255^2 < 2^24 and
    # |dot| < 2^24), so a dropped, doubled or mis-indexed term shows up.
    rng = np.random.default_rng(d * 100 + W)
    x = rint(rng, (20, d))
    y = rint(rng, (20, d))
    exact = ((x.astype(np.int64) - y.astype(np.int64)) ** 2).sum(1)
    assert np.array_equal(o.l2_simd_batch(x, y, W), exact.astype(np.float32))
    a = rint(rng, (20, d), -15, 16)
    b = rint(rng, (20, d), -15, 16)
    ip = (a.astype(np.int64) * b.astype(np.int64)).sum(1)
    assert np.array_equal(o.ip_simd_batch(a, b, W), ip.astype(np.float32))


@pytest.mark.parametrize("d", [1, 3, 7, 13, 15, 27, 31, 33, 127, 4, 8, 16, 32, 64])
def test_distances_match_float64_like_reference_test(d):
    # tests/test_distances_simd.mojo:25-65 (atol 1e-4 against a scalar loop).
    rng = np.random.default_rng(d)
    x = (rng.random(d) * 20 - 10).astype(np.float32)
    y = (rng.random(d) * 20 - 10).astype(np.float32)
    l2 = float(((x.astype(np.float64) - y) ** 2).sum())
    ip = float((x.astype(np.float64) * y).sum())
    assert abs(o.l2_simd(x, y, 4) - l2) <= 1e-4 * max(1.0, abs(l2))
    assert abs(o.ip_simd(x, y, 4) - ip) <= 1e-4 * max(1.0, abs(ip))


def test_sub_simd_accumulators_are_zero():
    # tests/test_distances_simd.mojo:68-84
    x = np.array([1, 2, 3, 4], np.float32)
    y = np.arange(4, dtype=np.float32) * 0.25
    assert abs(o.l2_simd(x, y, 8) - float(((x - y) ** 2).sum())) < 1e-6
    assert abs(o.ip_simd(x, y, 8) - float((x * y).sum())) < 1e-6


def _tree(v):
    v = list(v)
    h = len(v) // 2
    while h >= 1:
        v = [np.float32(v[j] + v[j + h]) for j in range(h)]
        h //= 2
    return v[0]


def _l2_w_order(x, y, W):
    # Independent numpy statement of the documented order (four W-lane FMA
    # accumulators, lane-wise ((a0+a1)+a2)+a3, halving tree). This pins the
    # C code to the documented order; the order itself is the restatement's
    # choice for Mojo's SIMD.reduce_add (parity unpinned against Mojo).
    x = x.astype(np.float32)
    y = y.astype(np.float32)
    d = x.size
    acc = [np.zeros(W, np.float32) for _ in range(4)]
    i = 0
    while i <= d - 4 * W:
        for t in range(4):
            df = (x[i + t * W:i + (t + 1) * W] - y[i + t * W:i + (t + 1) * W]).astype(np.float64)
            acc[t] = (df * df + acc[t].astype(np.float64)).astype(np.float32)  # exact fma
        i += 4 * W
    v = ((acc[0] + acc[1]) + acc[2]) + acc[3]
    while i <= d - W:
        df = (x[i:i + W] - y[i:i + W]).astype(np.float64)
        v = (df * df + v.astype(np.float64)).astype(np.float32)
        i += W
    assert i == d
    return _tree(v)


@pytest.mark.parametrize("W,d", [(8, 128), (8, 96), (64, 768), (64, 128), (8, 8)])
def test_documented_reduction_order(W, d):
    rng = np.random.default_rng(W + d)
    x = rng.standard_normal((8, d)).astype(np.float32)
    y = rng.standard_normal((8, d)).astype(np.float32)
    got = o.l2_simd_batch(x, y, W)
    want = np.array([_l2_w_order(x[i], y[i], W) for i in range(8)], np.float32)
    assert np.array_equal(got, want)


def test_short4_tree():
    x = np.array([1e8, 1.0, -1e8, 1.0], np.float32)
    y = np.zeros(4, np.float32)
    # (r0 + r2) + (r1 + r3) on squares: 1e16+1e16 then +2
    r = (x.astype(np.float32) ** 2)
    want = np.float32(np.float32(r[0] + r[2]) + np.float32(r[1] + r[3]))
    assert o.l2_short4(x, y) == want


def test_int8_and_uint8_dots_match_reference_patterns():
    for name, fn in (("i8_dot.npz", o.dot_i8), ("u8_dot.npz", o.dot_u8)):
        g = np.load(os.path.join(GOLDEN, name))
        off = 0
        for d, e in zip(g["dims"], g["expected"]):
            assert fn(g["x"][off:off + d], g["y"][off:off + d]) == e
            off += d


def test_l2_from_dot_clamps_at_zero():
    assert o.l2_from_dot(1, 1, 5) == 0
    assert o.l2_from_dot(100, 50, 20) == 110


def test_uint8_rebase_identity():
    # sum u_q u_b = sum s_q s_b + 128 (sum s_q + sum s_b) + 16384 d, s = u - 128
    rng = np.random.default_rng(3)
    for d in (1, 5, 64, 768):
        u = rng.integers(0, 256, (2, d)).astype(np.uint8)
        s = (u.astype(np.int64) - 128).astype(np.int8)
        lhs = o.dot_u8(u[0], u[1])
        rhs = o.dot_i8(s[0], s[1]) + 128 * (int(s[0].astype(np.int64).sum()) + int(s[1].astype(np.int64).sum())) + 16384 * d
        assert lhs == rhs


# ---------------------------------------------------------------- top-k
def test_reference_heap_is_not_tie_stable_but_canonical_is():
    d, i = o.heap_topk_reference(np.array([5, 5, 1], np.float32), 2)
    assert sorted(i.tolist()) == [1, 2]
    d, i = o.topk_canonical(np.array([5, 5, 1], np.float32), 2)
    assert i.tolist() == [2, 0]


def test_canonical_topk_matches_lexsort_and_reference_distance_multiset():
    rng = np.random.default_rng(1)
    for trial in range(200):
        n = int(rng.integers(1, 60))
        k = int(rng.integers(1, 20))
        dist = rng.integers(0, 7, n).astype(np.float32)
        cd, ci = o.topk_canonical(dist, k)
        order = np.lexsort((np.arange(n), dist))[:k]
        assert ci.tolist() == order.tolist()
        rd, ri = o.heap_topk_reference(dist, k)
        assert np.array_equal(np.sort(rd), np.sort(cd))


# ---------------------------------------------------------------- flat
def test_flat_known_answers():
    g = json.load(open(os.path.join(GOLDEN, "flat_known.json")))
    for key, metric in (("l2", o.L2), ("ip", o.IP)):
        c = g[key]
        D, I = o.flat_search(np.array(c["base"], np.float32), np.array([c["query"]], np.float32),
                             c["k"], metric)
        assert I[0].tolist() == c["expected_ids"]
        assert np.allclose(D[0], c["expected_dist"], atol=c["atol"])


def test_flat_canonical_ties_and_exclusion():
    g = np.load(os.path.join(GOLDEN, "ties.npz"))
    k = int(g["k"])
    D, I = o.flat_search(g["base"], g["q"], k)
    assert np.array_equal(I, g["ids"])
    assert np.array_equal(D, g["dist"])
    ex = np.zeros(g["base"].shape[0], np.uint8)
    ex[g["ids"][0, 0]] = 1
    D2, I2 = o.flat_search(g["base"], g["q"][:1], k, exclude=ex)
    assert g["ids"][0, 0] not in I2[0]
    assert I2[0].tolist() == [i for i in g["ids"][0].tolist() if i != g["ids"][0, 0]][: k - 1] + [I2[0, -1]]


def test_flat_pads_when_k_exceeds_n():
    D, I = o.flat_search(np.ones((3, 4), np.float32), np.zeros((1, 4), np.float32), 5)
    assert I[0, 3:].tolist() == [-1, -1] and D[0, 3:].tolist() == [0.0, 0.0]


# ---------------------------------------------------------------- k-means
def _lloyd_step64(x, c):
    d2 = ((x[:, None, :].astype(np.float64) - c[None, :, :]) ** 2).sum(-1)
    a = d2.argmin(1)
    out = c.copy()
    for j in range(c.shape[0]):
        m = a == j
        if m.any():
            s = x[m].astype(np.float64).sum(0)
            out[j] = (s.astype(np.float32) * np.float32(np.float32(1.0) / np.float32(m.sum())))
    return a, out


def test_kmeans_iteration_matches_lloyd_closed_form():
    rng = np.random.default_rng(5)
    x = rint(rng, (500, 12), 0, 50)
    init = x[rng.choice(500, 7, replace=False)].copy()
    for it in (1, 2, 3):
        c, a = o.kmeans_train(x, 7, it, init=init, sample_offset=0)
        ref_c = init.copy()
        for _ in range(it):
            ref_a, ref_c = _lloyd_step64(x, ref_c)
        # integer data: exact distances and exact sums, so bitwise equal
        assert np.array_equal(c, ref_c)


def test_kmeans_fp64_and_reference_fp32_sums_agree_on_integer_data():
    rng = np.random.default_rng(6)
    x = rint(rng, (2000, 16))
    init, off = o.kmeans_init(x, 20)
    c64, a64 = o.kmeans_train(x, 20, 5, init=init, sample_offset=off, fp64_sums=True)
    c32, a32 = o.kmeans_train(x, 20, 5, init=init, sample_offset=off, fp64_sums=False)
    assert np.array_equal(c64, c32) and np.array_equal(a64, a32)


def test_kmeans_empty_cluster_reseed_rule():
    x = np.arange(40, dtype=np.float32).reshape(20, 2)
    init = np.array([[0, 1], [1e6, 1e6]], np.float32)  # cluster 1 stays empty
    c, a = o.kmeans_train(x, 2, 1, init=init, sample_offset=3)
    src = (3 + 1 * 2 + 1) % 20
    assert np.array_equal(c[1], x[src])
    assert (a == 0).all()


def test_kmeans_identical_points_k_gt_n_and_reproducible():
    # tests/test_kmeans.mojo:22-72
    c, _ = o.kmeans_train(np.ones((100, 16), np.float32), 5, 5)
    assert not np.isnan(c).any()
    c, _ = o.kmeans_train(np.arange(80, dtype=np.float32).reshape(5, 16), 10, 5)
    assert c.shape == (10, 16)
    i = np.arange(257 * 8)
    x = (((i * 37 + i // 8 * 11) % 101) / 50.0 - 1.0).astype(np.float32).reshape(257, 8)
    c1, _ = o.kmeans_train(x, 17, 5)
    c2, _ = o.kmeans_train(x, 17, 5)
    assert np.array_equal(c1, c2)


def test_philox_known_answer():
    # Philox-4x32-10 known-answer vectors (Random123 kat_vectors):
    # counter 0, key 0 -> 6627e8d5 e169c58d bc57ac4c 9b00dbd8
    out = np.zeros(4, np.uint32)
    o.lib().mvref_philox4x32_10(0, 0, out)
    assert [hex(v) for v in out] == ["0x6627e8d5", "0xe169c58d", "0xbc57ac4c", "0x9b00dbd8"]


# ---------------------------------------------------------------- PQ
def _uniform(seed, n, d):
    return (np.random.default_rng(seed).random((n, d)) * 2 - 1).astype(np.float32)


def test_pq_reconstruction_mse_and_adc_identity():
    # tests/test_pq.mojo:8-114
    x = _uniform(181, 1000, 16)
    cb = o.pq_train(x, 4)
    codes = o.pq_encode(x, cb)
    dec = o.pq_decode(codes, cb)
    mse = float(((x.astype(np.float64) - dec) ** 2).sum(1).mean())
    assert mse < 1.0
    lut = o.pq_lut(x[0], cb)
    adc = np.float32(0)
    for m in range(4):
        adc = np.float32(adc + lut[m, codes[0, m]])
    assert abs(float(adc) - float(((x[0].astype(np.float64) - dec[0]) ** 2).sum())) < 1e-5


def test_pq_encode_is_bruteforce_argmin():
    x = _uniform(3, 400, 32)
    cb = o.pq_train(x, 4, niter=3)
    codes = o.pq_encode(x, cb)
    for m in range(4):
        sub = x[:, m * 8:(m + 1) * 8].astype(np.float64)
        d2 = ((sub[:, None, :] - cb[m][None, :, :]) ** 2).sum(-1)
        best = d2.argmin(1)
        srt = np.sort(d2, 1)
        clear = (srt[:, 1] - srt[:, 0]) > 1e-4
        assert np.array_equal(codes[clear, m], best[clear])


def test_pq_lut_l2_and_ip():
    x = _uniform(4, 300, 24)
    cb = o.pq_train(x, 3, niter=2)
    q = _uniform(5, 1, 24)[0]
    l2 = o.pq_lut(q, cb, o.L2)
    ip = o.pq_lut(q, cb, o.IP)
    for m in range(3):
        sub = q[m * 8:(m + 1) * 8].astype(np.float64)
        assert np.allclose(l2[m], ((cb[m] - sub) ** 2).sum(1), rtol=1e-5, atol=1e-6)
        assert np.allclose(ip[m], -(cb[m] @ sub), rtol=1e-5, atol=1e-6)


# ---------------------------------------------------------------- SQ8
def _round_half_away(v):
    return np.sign(v) * np.floor(np.abs(v) + 0.5)


def test_sq8_encode_formula_and_norms():
    rng = np.random.default_rng(8)
    x = (rng.standard_normal(777) * 3).astype(np.float32)
    mn, inv = np.float32(x.min()), np.float32(1.0) / np.float32((x.max() - x.min()) / np.float32(255.0))
    code, norm = o.sq8_encode(x, False, float(mn), float(inv))
    want = np.clip(_round_half_away(((x - mn) * inv).astype(np.float32)), 0, 255).astype(np.uint8)
    assert np.array_equal(code, want)
    assert norm == int((want.astype(np.int64) ** 2).sum())
    inv2 = np.float32(1.0) / np.float32(np.abs(x).max() / np.float32(127.0))
    code, norm = o.sq8_encode(x, True, 0.0, float(inv2))
    want = np.clip(_round_half_away((x * inv2).astype(np.float32)), -127, 127).astype(np.int8)
    assert np.array_equal(code, want)


def test_sq8_integer_data_codes_are_the_data_and_search_is_exact():
    rng = np.random.default_rng(9)
    x = rint(rng, (3000, 32))
    x[0, 0], x[1, 0] = 0, 255
    ix = o.SQ8Index(32)
    ix.add(x)
    gmin, gmax, scale, n = ix.params()
    assert (gmin, gmax, scale, n) == (0.0, 255.0, 1.0, 3000)
    codes, norms = ix.codes()
    assert np.array_equal(codes, x.astype(np.uint8))
    q = rint(rng, (30, 32))
    D, I = ix.search(q, 10)
    Df, If = o.flat_search(x, q, 10)
    assert np.array_equal(I, If) and np.array_equal(D, Df)


def test_sq8_ip_ordering_and_requantize_scale():
    # tests/test_metrics.mojo:56-93
    ix = o.SQ8Index(3, o.IP)
    ix.add(np.array([[1, 0, 0], [-1, 0, 0], [0.5, 0, 0]], np.float32))
    D, I = ix.search(np.array([[1, 0, 0]], np.float32), 3)
    assert I[0].tolist() == [0, 2, 1] and D[0, 0] > D[0, 1] > D[0, 2]
    ix = o.SQ8Index(2, o.IP)
    ix.add(np.array([[1.0, 0.25]], np.float32))
    ix.add(np.array([[-4.0, 0.0]], np.float32))
    assert abs(ix.params()[2] - 4.0 / 127.0) < 1e-6


def _uniform_dataset():
    g = np.random.default_rng(42)
    return (g.random((5000, 32)) * 2 - 1).astype(np.float32), (g.random((100, 32)) * 2 - 1).astype(np.float32)


def _recall(gt, got):
    return np.mean([len(set(a) & set(b)) / len(a) for a, b in zip(gt.tolist(), got.tolist())])


def test_sq8_ip_recall_and_rerank():
    # tests/test_accuracy.mojo:190-230 (recall >= 0.95 vs exact IP)
    db, q = _uniform_dataset()
    _, gt = o.flat_search(db, q, 10, o.IP)
    ix = o.SQ8Index(32, o.IP)
    ix.add(db)
    _, I = ix.search(q, 10)
    assert _recall(gt, I) >= 0.95
    _, I4 = ix.search(q, 10, k_factor=4)
    assert _recall(gt, I4) >= _recall(gt, I)


# ---------------------------------------------------------------- IVF-PQ
def test_ivfpq_accuracy_threshold():
    # tests/test_accuracy.mojo:129-161: recall@10 >= 0.50 at nlist 64, M 8, nprobe 16
    db, q = _uniform_dataset()
    _, gt = o.flat_search(db, q, 10)
    ix = o.IVFPQIndex(32, 64, 8)
    ix.train(db)
    ix.add(db)
    _, I = ix.search(q, 10, 16)
    assert _recall(gt, I) >= 0.50
    _, Ir = ix.search(q, 10, 16, k_factor=8)
    assert _recall(gt, Ir) > _recall(gt, I)


def test_ivfpq_lifecycle_self_query():
    # tests/test_ivf_pq.mojo:48-83: 256 x 8 closed form, nlist 8, M 2, nprobe 4
    rows = np.load(os.path.join(GOLDEN, "ivfpq_rows.npz"))["rows"]
    ix = o.IVFPQIndex(8, 8, 2)
    ix.train(rows)
    ix.add(rows)
    D, I = ix.search(rows[:1], 5, 4)
    assert I[0, 0] == 0 and D[0, 0] < 1e-4


def test_ivfpq_exhaustive_equals_adc_bruteforce():
    rng = np.random.default_rng(11)
    x = rint(rng, (2000, 16), 0, 64)
    ix = o.IVFPQIndex(16, 8, 4)
    ix.train(x, niter_coarse=5, niter_pq=5)
    ix.add(x, keep_vectors=True)
    coarse, pq = ix.codebooks()
    q = rint(rng, (10, 16), 0, 64)
    D, I = ix.search(q, 20, 8)  # nprobe = nlist: every list
    # reconstruct each vector as c_list + decode(code) from the stored lists
    recon = np.zeros((2000, 16), np.float64)
    for l in range(8):
        ids, codes = ix.list(l)
        assert np.all(np.diff(ids) > 0)  # append order == id order
        recon[ids] = coarse[l] + o.pq_decode(codes, pq)
    d2 = ((q[:, None, :].astype(np.float64) - recon[None]) ** 2).sum(-1)
    for i in range(10):
        got = d2[i, I[i]]
        assert np.allclose(D[i], got, rtol=1e-4, atol=1e-3)
        kth = np.sort(d2[i])[19]
        assert (got <= kth + 1e-3).all()


def test_ivfpq_reference_threading_variant_gives_the_same_results():
    # the CPU-baseline variant with the reference's threading (coarse
    # parallel, scan serial) is the same search
    rng = np.random.default_rng(13)
    x = rint(rng, (3000, 16), 0, 64)
    ix = o.IVFPQIndex(16, 12, 4)
    ix.train(x, niter_coarse=3, niter_pq=3)
    ix.add(x)
    q = rint(rng, (40, 16), 0, 64)
    for kf in (1, 3):
        D1, I1 = ix.search(q, 7, 5, k_factor=kf)
        D2, I2 = ix.search_refstruct(q, 7, 5, k_factor=kf)
        assert np.array_equal(I1, I2) and np.array_equal(D1, D2)


def test_ivfpq_padding_exclusion_and_untrained():
    ix = o.IVFPQIndex(8, 4, 2)
    D, I = ix.search(np.zeros((2, 8), np.float32), 3, 2)
    assert (I == -1).all() and (D == np.float32(1e38)).all()
    rows = np.load(os.path.join(GOLDEN, "ivfpq_rows.npz"))["rows"]
    ix.train(rows)
    ix.add(rows)
    ex = np.zeros(256, np.uint8)
    ex[0] = 1
    _, I = ix.search(rows[:1], 1, 4, exclude=ex)
    assert I[0, 0] != 0
    D, I = ix.search(rows[:1], 300, 1)
    n_found = (I[0] >= 0).sum()
    assert n_found < 300 and (D[0, n_found:] == np.float32(1e38)).all()


def test_ivfpq_rerank_with_every_candidate_is_exact_search():
    rng = np.random.default_rng(12)
    x = rint(rng, (600, 16), 0, 32)
    ix = o.IVFPQIndex(16, 4, 4)
    ix.train(x, niter_coarse=3, niter_pq=3)
    ix.add(x)
    q = rint(rng, (5, 16), 0, 32)
    D, I = ix.search(q, 5, 4, k_factor=120)  # k * k_factor = every row
    Df, If = o.flat_search(x, q, 5)
    assert np.array_equal(I, If) and np.array_equal(D, Df)


def _integer_ivfpq(seed, d, nlist, M, metric, n=1500):
    """An IVF-PQ index whose codebooks are small integers, so that on integer
    rows and queries every FP32 sum of the search is exact: the oracle must
    then equal float64 arithmetic on the reconstructed rows bit for bit."""
    rng = np.random.default_rng(seed)
    coarse = rint(rng, (nlist, d), -6, 7)
    pq = rint(rng, (M, 256, d // M), -3, 4)
    ix = o.IVFPQIndex(d, nlist, M, metric)
    ix.set_codebooks(coarse, pq)
    x = rint(rng, (n, d), -8, 9)
    ix.add(x)
    owner = np.full(n, -1)
    recon = np.zeros((n, d), np.float64)
    for l in range(nlist):
        ids, codes = ix.list(l)
        owner[ids] = l
        recon[ids] = coarse[l].astype(np.float64) + o.pq_decode(codes, pq)
    assert (owner >= 0).all()
    return ix, rng, x, coarse, owner, recon


@pytest.mark.parametrize("M", [4, 2])  # dsub 4 (short form) and 8 (W = 4 LUT)
def test_ivfpq_inner_product_is_minus_dot_with_reconstruction(M):
    # index_ivf_pq.mojo:288-293,316-324: an inner-product search scores an
    # entry of list l as -<q, c_l> + sum_m LUT_IP(q)[m][code_m], i.e.
    # -<q, c_l + decode(code)>, and returns it un-negated (:357-360). Every
    # sum is exact here, so a dropped coarse term, a wrong sign or a
    # transposed LUT index changes the result; exact ties order by id.
    d, nlist = 16, 8
    ix, rng, x, coarse, owner, recon = _integer_ivfpq(31 + M, d, nlist, M, o.IP)
    q = rint(rng, (15, d), -8, 9)
    q64 = q.astype(np.float64)
    ids_all = np.arange(x.shape[0], dtype=np.int64)
    k = 40
    D, I = ix.search(q, k, nlist)  # every list
    for i in range(q.shape[0]):
        ed, ei = _canonical_topk_f64(-(recon @ q64[i]), ids_all, k)
        assert np.array_equal(I[i], ei)
        assert np.array_equal(D[i].astype(np.float64), -ed)
    nprobe = 3
    D, I = ix.search(q, k, nprobe)
    for i in range(q.shape[0]):
        probed = np.lexsort((np.arange(nlist), -(coarse.astype(np.float64) @ q64[i])))[:nprobe]
        rows = np.flatnonzero(np.isin(owner, probed)).astype(np.int64)
        ed, ei = _canonical_topk_f64(-(recon[rows] @ q64[i]), rows, k)
        assert np.array_equal(I[i], ei)
        assert np.array_equal(D[i].astype(np.float64), -ed)


@pytest.mark.parametrize("metric", [o.L2, o.IP])
def test_ivfpq_rerank_reorders_exactly_the_adc_candidates(metric):
    # k_factor: the k * k_factor ADC candidates of the search, re-scored with
    # exact distances to the stored rows (index_hnsw.mojo:449-482), best k by
    # (distance, id). Integer data: the exact distances are float64's.
    d, nlist, M = 16, 8, 4
    ix, rng, x, coarse, owner, recon = _integer_ivfpq(41, d, nlist, M, metric)
    q = rint(rng, (12, d), -8, 9)
    x64, q64 = x.astype(np.float64), q.astype(np.float64)
    k, kf, nprobe = 7, 6, 4
    _, cand = ix.search(q, k * kf, nprobe)
    D, I = ix.search(q, k, nprobe, k_factor=kf)
    for i in range(q.shape[0]):
        c = cand[i][cand[i] >= 0]
        ex = ((x64[c] - q64[i]) ** 2).sum(-1) if metric == o.L2 else -(x64[c] @ q64[i])
        ed, ei = _canonical_topk_f64(ex, c, k)
        assert np.array_equal(I[i], ei)
        assert np.array_equal(D[i].astype(np.float64), ed if metric == o.L2 else -ed)
    # every stored row as a candidate: the exact search over the probed lists
    D, I = ix.search(q, k, nlist, k_factor=x.shape[0])
    Df, If = o.flat_search(x, q, k, metric)
    assert np.array_equal(I, If) and np.array_equal(D, Df)


@pytest.mark.parametrize("metric", [o.L2, o.IP])
def test_ivfpq_add_with_ids_stores_and_returns_the_callers_ids(metric):
    # index_ivf_pq.mojo:150-213: the lists hold the given ids in input order;
    # a search returns them, the exclusion mask is indexed by them, and the
    # rerank scores the row added under each id. Against the same index
    # built with sequential ids, every result id maps through the id table.
    rng = np.random.default_rng(71 + metric)
    x = rint(rng, (900, 16), -8, 9)
    q = rint(rng, (10, 16), -8, 9)
    coarse = rint(rng, (6, 16), -6, 7)
    pq = rint(rng, (4, 256, 4), -3, 4)
    ids = rng.permutation(100000)[:900].astype(np.int64)
    a = o.IVFPQIndex(16, 6, 4, metric)
    b = o.IVFPQIndex(16, 6, 4, metric)
    for ix in (a, b):
        ix.set_codebooks(coarse, pq)
    a.add_with_ids(x[:500], ids[:500])
    a.add_with_ids(x[500:], ids[500:])
    b.add(x)
    for l in range(6):
        ia, ca = a.list(l)
        ib, cb = b.list(l)
        assert np.array_equal(ia, ids[ib]) and np.array_equal(ca, cb)
    ex_rows = np.zeros(900, np.uint8)
    ex_rows[::7] = 1
    ex_ids = np.zeros(100000, np.uint8)
    ex_ids[ids[ex_rows == 1]] = 1
    # every probed, unexcluded entry with its ADC distance, from the
    # sequential-id index, then the canonical order under the custom ids
    Dall, Rall = b.search(q, 900, 3, exclude=ex_rows)
    x64, q64 = x.astype(np.float64), q.astype(np.float64)
    k, kf = 8, 5
    Da, Ia = a.search(q, k, 3, exclude=ex_ids)
    Dr, Ir = a.search(q, k, 3, k_factor=kf, exclude=ex_ids)
    for i in range(q.shape[0]):
        ok = Rall[i] >= 0
        adc, cid, rows = Dall[i][ok].astype(np.float64), ids[Rall[i][ok]], Rall[i][ok]
        sgn = 1.0 if metric == o.L2 else -1.0  # IP results are un-negated
        order = np.lexsort((cid, sgn * adc))
        assert np.array_equal(Ia[i], cid[order[:k]])
        assert np.array_equal(Da[i].astype(np.float64), adc[order[:k]])
        cand = order[:k * kf]  # the rerank's candidates, scored on the row of each id
        ex = (((x64[rows[cand]] - q64[i]) ** 2).sum(-1) if metric == o.L2
              else -(x64[rows[cand]] @ q64[i]))
        o2 = np.lexsort((cid[cand], ex))[:k]
        assert np.array_equal(Ir[i], cid[cand][o2])
        assert np.array_equal(Dr[i].astype(np.float64), sgn * ex[o2])


def _ip_w_order(x, y, W):
    """inner_product_simd[W]'s documented order, as _l2_w_order for L2."""
    x = x.astype(np.float32)
    y = y.astype(np.float32)
    d = x.size
    acc = [np.zeros(W, np.float32) for _ in range(4)]
    i = 0
    while i <= d - 4 * W:
        for t in range(4):
            p = x[i + t * W:i + (t + 1) * W].astype(np.float64) * y[i + t * W:i + (t + 1) * W]
            acc[t] = (p + acc[t].astype(np.float64)).astype(np.float32)
        i += 4 * W
    v = ((acc[0] + acc[1]) + acc[2]) + acc[3]
    while i <= d - W:
        p = x[i:i + W].astype(np.float64) * y[i:i + W]
        v = (p + v.astype(np.float64)).astype(np.float32)
        i += W
    assert i == d
    return _tree(v)


@pytest.mark.parametrize("metric", [o.L2, o.IP])
def test_sq8_rerank_uses_w64_exact_distances_of_the_code_candidates(metric):
    # index_hnsw.mojo:433-496: the SQ8 candidates are re-scored with
    # l2_distance_simd[64] / -inner_product_simd[64] against the FP32 rows and
    # the best k kept by (distance, id). The W = 64 distances come from the
    # independent numpy statement of that order (float rows, d = 128: two
    # 64-wide blocks and the halving tree).
    rng = np.random.default_rng(51 + metric)
    d, n = 128, 300
    x = rng.standard_normal((n, d)).astype(np.float32)
    q = rng.standard_normal((6, d)).astype(np.float32)
    ix = o.SQ8Index(d, metric)
    ix.add(x)

    def exact(qi, rows):
        if metric == o.L2:
            return np.array([_l2_w_order(q[qi], x[r], 64) for r in rows], np.float64)
        return -np.array([_ip_w_order(q[qi], x[r], 64) for r in rows], np.float64)

    k, kf = 5, 4
    _, cand = ix.search(q, k * kf)
    D, I = ix.search(q, k, k_factor=kf)
    for i in range(q.shape[0]):
        ed, ei = _canonical_topk_f64(exact(i, cand[i]), cand[i], k)
        assert np.array_equal(I[i], ei)
        assert np.array_equal(D[i].astype(np.float64), ed if metric == o.L2 else -ed)
    # k * k_factor covering every row: the W = 64 exact search
    D, I = ix.search(q[:3], k, k_factor=n)
    for i in range(3):
        ed, ei = _canonical_topk_f64(exact(i, np.arange(n)), np.arange(n, dtype=np.int64), k)
        assert np.array_equal(I[i], ei)
        assert np.array_equal(D[i].astype(np.float64), ed if metric == o.L2 else -ed)


@pytest.mark.parametrize("metric", [o.L2, o.IP])
def test_sq8_full_rerank_on_integer_rows_is_flat_search(metric):
    # Integer rows with the full code range (scale 1: codes are the rows):
    # a rerank over every row equals the exact Flat search, ids and distances.
    rng = np.random.default_rng(61 + metric)
    lo, hi = (0, 256) if metric == o.L2 else (-127, 128)
    x = rint(rng, (400, 24), lo, hi)
    x[0, 0], x[1, 0] = lo, hi - 1
    q = rint(rng, (9, 24), lo, hi)
    ix = o.SQ8Index(24, metric)
    ix.add(x)
    assert ix.params()[2] == 1.0
    D, I = ix.search(q, 12, k_factor=400)
    Df, If = o.flat_search(x, q, 12, metric)
    assert np.array_equal(I, If) and np.array_equal(D, Df)


# ---------------------------------------------------------------- IVF-Flat
def _canonical_topk_f64(d2, ids, k):
    """k smallest (distance, id) pairs in lexicographic order, from float64."""
    order = np.lexsort((ids, d2))[:k]
    return d2[order], ids[order]


def test_ivfflat_accuracy_threshold():
    # tests/test_accuracy.mojo:96-126: recall@10 >= 0.75 at nlist 64, nprobe 16
    db, q = _uniform_dataset()
    _, gt = o.flat_search(db, q, 10)
    ix = o.IVFFlatIndex(32, 64)
    ix.train(db)
    ix.add(db)
    _, I = ix.search(q, 10, 16)
    assert _recall(gt, I) >= 0.75
    _, Iall = ix.search(q, 10, 64)
    assert _recall(gt, Iall) == 1.0


@pytest.mark.parametrize("metric", [o.L2, o.IP])
def test_ivfflat_exhaustive_probe_is_exact_search(metric):
    # nprobe = nlist scans every row: on integer data every summation order is
    # exact, so the result is the float64 brute force with (distance, id) ties.
    rng = np.random.default_rng(21)
    x = rint(rng, (1500, 24), 0, 8)  # few distinct values: many exact ties
    q = rint(rng, (12, 24), 0, 8)
    ix = o.IVFFlatIndex(24, 6, metric)
    ix.train(x, niter=4)
    ids = rng.permutation(100000)[:1500].astype(np.int64)  # custom, unordered ids
    ix.add(x, ids=ids)
    D, I = ix.search(q, 30, 6)
    x64, q64 = x.astype(np.float64), q.astype(np.float64)
    for i in range(12):
        dd = ((q64[i] - x64) ** 2).sum(-1) if metric == o.L2 else -(x64 @ q64[i])
        ed, ei = _canonical_topk_f64(dd, ids, 30)
        assert np.array_equal(I[i], ei)
        assert np.array_equal(D[i].astype(np.float64), ed if metric == o.L2 else -ed)


@pytest.mark.parametrize("metric", [o.L2, o.IP])
def test_ivfflat_probes_only_the_nearest_lists(metric):
    # nprobe < nlist: the result is the brute force over the rows whose list
    # is among the query's nprobe nearest centroids (canonical ties), and each
    # row sits in the list of its own nearest centroid.
    rng = np.random.default_rng(22)
    x = rint(rng, (3000, 16), 0, 16)
    q = rint(rng, (20, 16), 0, 16)
    nlist, nprobe = 12, 3
    ix = o.IVFFlatIndex(16, nlist, metric)
    ix.train(x, niter=5)
    ix.add(x)
    cen = ix.centroids().astype(np.float64)
    x64, q64 = x.astype(np.float64), q.astype(np.float64)
    cd = (((x64[:, None] - cen[None]) ** 2).sum(-1) if metric == o.L2 else -(x64 @ cen.T))
    owner = np.array([np.lexsort((np.arange(nlist), r))[0] for r in cd])
    for l in range(nlist):
        lid, rows = ix.list(l)
        assert np.array_equal(lid, np.flatnonzero(owner == l))  # append order
        assert np.array_equal(rows, x[lid])
    D, I = ix.search(q, 10, nprobe)
    for i in range(20):
        qc = ((q64[i] - cen) ** 2).sum(-1) if metric == o.L2 else -(cen @ q64[i])
        probed = np.lexsort((np.arange(nlist), qc))[:nprobe]
        rows = np.flatnonzero(np.isin(owner, probed))
        dd = ((q64[i] - x64[rows]) ** 2).sum(-1) if metric == o.L2 else -(x64[rows] @ q64[i])
        ed, ei = _canonical_topk_f64(dd, rows.astype(np.int64), 10)
        assert np.array_equal(I[i], ei)
        assert np.array_equal(D[i].astype(np.float64), ed if metric == o.L2 else -ed)


def test_ivfflat_reference_known_answers():
    # tests/test_ivf_flat.mojo:69-88: one vector, nlist 1 -> id 0 at distance 0
    data = np.arange(4, dtype=np.float32)[None]
    ix = o.IVFFlatIndex(4, 1)
    ix.train(data)
    ix.add(data)
    D, I = ix.search(data, 1, 1)
    assert I[0, 0] == 0 and D[0, 0] == 0.0
    # :91-137: k = 5 over one stored row with id 42 -> tail (1e38, -1)
    zero = np.zeros((1, 4), np.float32)
    ix = o.IVFFlatIndex(4, 1)
    ix.train(zero)
    ix.add(zero, ids=np.array([42]))
    D, I = ix.search(zero, 5, 1)
    assert I[0].tolist() == [42, -1, -1, -1, -1]
    assert D[0, 0] == 0.0 and (D[0, 1:] == np.float32(1e38)).all()
    # :9-66: custom ids i * 10, nprobe 3 of 10 lists; every result is a stored id
    rng = np.random.default_rng(101)
    x = (rng.random((1000, 16)) * 2 - 1).astype(np.float32)
    ix = o.IVFFlatIndex(16, 10)
    ix.train(x)
    ix.add(x, ids=np.arange(1000) * 10)
    D, I = ix.search(x[:10], 5, 3)
    assert (I % 10 == 0).all() and (I[:, 0] == np.arange(10) * 10).all()
    assert (np.diff(D, axis=1) >= 0).all()


def test_ivfflat_untrained_empty_oversized_nprobe_and_exclusion():
    ix = o.IVFFlatIndex(8, 4)
    rows = np.load(os.path.join(GOLDEN, "ivfpq_rows.npz"))["rows"]
    ix.add(rows)  # dropped: not trained (index_ivf_flat.mojo:102-104)
    D, I = ix.search(rows[:2], 3, 2)
    assert (I == -1).all() and (D == np.float32(1e38)).all()
    ix.train(rows)
    D, I = ix.search(rows[:2], 3, 2)  # trained, empty
    assert (I == -1).all() and (D == np.float32(1e38)).all()
    ix.add(rows)
    D1, I1 = ix.search(rows[:3], 4, 4)
    D2, I2 = ix.search(rows[:3], 4, 400)  # nprobe clamps to nlist
    assert np.array_equal(I1, I2) and np.array_equal(D1, D2)
    ex = np.zeros(256, np.uint8)
    ex[I1[:, 0]] = 1
    _, Ie = ix.search(rows[:3], 4, 4, exclude=ex)
    assert np.array_equal(Ie[:, :3], I1[:, 1:])


# ---- distance computers for a graph search (HNSW-facing batch service) ----
# The two FP32 summation orders are pinned against an exact rational
# emulation written from the reference text: every binary32 operation is
# rounded exactly (round to nearest, ties to even) from its exact rational
# value, fma rounding once.
from fractions import Fraction


def _ex_r32(v):
    """Exact binary32 rounding of a rational (normal range)."""
    if v == 0:
        return Fraction(0)
    sign = -1 if v < 0 else 1
    a = abs(v)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    while Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    m = a * Fraction(2) ** (23 - e)
    q, r = divmod(m.numerator, m.denominator)
    if 2 * r > m.denominator or (2 * r == m.denominator and q % 2 == 1):
        q += 1
    return sign * Fraction(q) * Fraction(2) ** (e - 23)


def _ex_tree(v):  # SIMD.reduce_add as the halving tree
    v = list(v)
    h = len(v) // 2
    while h >= 1:
        for j in range(h):
            v[j] = _ex_r32(v[j] + v[j + h])
        h //= 2
    return v[0]


def _ex_term(a, b, ip):
    return _ex_r32(a * b) if ip else _ex_r32(_ex_r32(a - b) ** 2)


def _ex_fma_term(a, b, acc, ip):
    return _ex_r32(a * b + acc) if ip else _ex_r32(_ex_r32(a - b) ** 2 + acc)


def _ex_emul_single8(x, y, ip):
    """l2_distance_simd[8] / inner_product_simd[8] (distances.mojo:43-228)."""
    d = len(x)
    if d < 8:
        r = Fraction(0)
        for i in range(d):
            r = _ex_r32(r + _ex_term(x[i], y[i], ip))
        return r
    acc = [[Fraction(0)] * 8 for _ in range(4)]
    i = 0
    while i <= d - 32:
        for b in range(4):
            for j in range(8):
                acc[b][j] = _ex_fma_term(x[i + 8 * b + j], y[i + 8 * b + j], acc[b][j], ip)
        i += 32
    v = [_ex_r32(_ex_r32(_ex_r32(acc[0][j] + acc[1][j]) + acc[2][j]) + acc[3][j]) for j in range(8)]
    while i <= d - 8:
        v = [_ex_fma_term(x[i + j], y[i + j], v[j], ip) for j in range(8)]
        i += 8
    res = _ex_tree(v)
    while i <= d - 4:
        res = _ex_r32(res + _ex_tree([_ex_term(x[i + j], y[i + j], ip) for j in range(4)]))
        i += 4
    while i < d:
        res = _ex_r32(res + _ex_term(x[i], y[i], ip))
        i += 1
    return res


def _ex_emul_batch8(x, y, ip):
    """One row of l2_distance_simd_batch4[8] / inner_product_simd_batch4[8]
    (distances.mojo:541-660)."""
    d = len(x)
    acc = [Fraction(0)] * 8
    i = 0
    while i + 8 <= d:
        acc = [_ex_fma_term(x[i + j], y[i + j], acc[j], ip) for j in range(8)]
        i += 8
    res = _ex_tree(acc)
    while i < d:
        res = _ex_r32(res + _ex_term(x[i], y[i], ip))
        i += 1
    return res


@pytest.mark.parametrize("d", [5, 8, 13, 32, 36, 45, 64, 71])
def test_distance_orders_match_exact_emulation(d):
    rng = np.random.default_rng(900 + d)
    differ = 0
    for trial in range(4):
        x = (rng.standard_normal(d) * rng.choice([1.0, 37.0])).astype(np.float32)
        y = (rng.standard_normal(d) * 3.0).astype(np.float32)
        xf = [Fraction(float(v)) for v in x]
        yf = [Fraction(float(v)) for v in y]
        for ip in (False, True):
            s_exp = float(_ex_emul_single8(xf, yf, ip))
            b_exp = float(_ex_emul_batch8(xf, yf, ip))
            s_got = o.ip_simd(x, y, 8) if ip else o.l2_simd(x, y, 8)
            b_got = o.ip_simd_b4(x, y, 8) if ip else o.l2_simd_b4(x, y, 8)
            assert np.float32(s_got) == np.float32(s_exp), (d, trial, ip, "single")
            assert np.float32(b_got) == np.float32(b_exp), (d, trial, ip, "batch4")
            differ += s_exp != b_exp
    if d >= 32:  # the two orders really differ once 32-element blocks exist
        assert differ > 0


def test_flat_distance_computer_batches():
    rng = np.random.default_rng(31)
    base = rng.standard_normal((300, 45)).astype(np.float32)
    q = rng.standard_normal((7, 45)).astype(np.float32)
    ids = rng.integers(0, 300, (7, 20))
    for metric in (o.L2, o.IP):
        for order in (0, 1):
            got = o.flat_distances(base, q, ids, metric, order)
            fn = (o.l2_simd if order == 0 else o.l2_simd_b4) if metric == o.L2 else \
                 (o.ip_simd if order == 0 else o.ip_simd_b4)
            for i in range(7):
                for j in range(20):
                    v = fn(q[i], base[ids[i, j]], 8) if order == 0 else fn(base[ids[i, j]], q[i], 8)
                    assert got[i, j] == np.float32(v if metric == o.L2 else -v)
    # integer rows: every order is exact, so both equal the float64 brute force
    bi = rng.integers(-20, 21, (50, 36)).astype(np.float32)
    qi = rng.integers(-20, 21, (3, 36)).astype(np.float32)
    idx = rng.integers(0, 50, (3, 9))
    exact_l2 = ((qi[:, None, :].astype(np.float64) - bi[idx]) ** 2).sum(-1)
    exact_ip = -(qi[:, None, :].astype(np.float64) * bi[idx]).sum(-1)
    for order in (0, 1):
        assert np.array_equal(o.flat_distances(bi, qi, idx, o.L2, order), exact_l2.astype(np.float32))
        assert np.array_equal(o.flat_distances(bi, qi, idx, o.IP, order), exact_ip.astype(np.float32))
    a, b = rng.integers(0, 50, 30), rng.integers(0, 50, 30)
    assert np.array_equal(o.flat_pair_distances(bi, a, b, o.L2),
                          ((bi[a].astype(np.float64) - bi[b]) ** 2).sum(-1).astype(np.float32))


def test_sq8_distance_computer_batches():
    rng = np.random.default_rng(32)
    x = rng.standard_normal((400, 24)).astype(np.float32)
    q = rng.standard_normal((5, 24)).astype(np.float32)
    ids = rng.integers(0, 400, (5, 30))
    for metric in (o.L2, o.IP):
        ix = o.SQ8Index(24, metric)
        ix.add(x)
        gmin, _, scale, _ = ix.params()
        codes, norms = ix.codes()
        got = ix.distances(q, ids)
        ssq = np.float32(scale) * np.float32(scale)
        for i in range(5):
            # the query codec of get_distance_computer (index_flat_sq8.mojo:536-555)
            qc, qn = o.sq8_encode_vec(q[i], metric != o.L2, 0.0 if metric != o.L2 else gmin,
                                      np.float32(1.0) / np.float32(scale))
            for j in range(30):
                c = codes[ids[i, j]]
                if metric == o.IP:
                    dot = int((qc.view(np.int8).astype(np.int64) * c.view(np.int8).astype(np.int64)).sum())
                    want = np.float32(-np.float32(dot)) * ssq
                else:
                    dot = int((qc.astype(np.int64) * c.astype(np.int64)).sum())
                    v = max(int(qn) + int(norms[ids[i, j]]) - 2 * dot, 0)
                    want = np.float32(v) * ssq
                assert got[i, j] == np.float32(want), (metric, i, j)
        a, b = rng.integers(0, 400, 40), rng.integers(0, 400, 40)
        gp = ix.pair_distances(a, b)
        for p in range(40):
            ca, cb = codes[a[p]], codes[b[p]]
            if metric == o.IP:
                dot = int((ca.view(np.int8).astype(np.int64) * cb.view(np.int8).astype(np.int64)).sum())
                want = np.float32(-np.float32(dot)) * ssq
            else:
                dot = int((ca.astype(np.int64) * cb.astype(np.int64)).sum())
                want = np.float32(max(int(norms[a[p]]) + int(norms[b[p]]) - 2 * dot, 0)) * ssq
            assert gp[p] == np.float32(want)
