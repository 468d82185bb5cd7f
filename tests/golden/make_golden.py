"""Writes the small known-answer fixtures under tests/golden/.

The inputs restate the closed-form data of the reference's own tests
(MojoVec v0.8.2); the expected outputs are computed here with exact integer
or float64 numpy arithmetic, independent of the oracle and of the GPU code:

  i8_dot.npz      tests/test_distances_simd.mojo:87-129 patterns, exact dots
  u8_dot.npz      the same patterns rebased to uint8 (x + 128), exact dots
  flat_known.json tests/test_index_flat.mojo:6-81 (L2 0.04 / 3.24, IP 12 / 8)
  ivfpq_rows.npz  tests/test_ivf_pq.mojo:30-38 256 x 8 dataset
  ties.npz        integer distances with deliberate exact ties and their
                  canonical (dist, id) top-k from numpy lexsort

Run: python tests/golden/make_golden.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def i8_patterns():
    dims = [1, 7, 16, 63, 64, 127, 128, 129]
    xs, ys, exp = [], [], []
    for d in dims:
        i = np.arange(d, dtype=np.int64)
        x = (i * 37) % 255 - 127
        y = (i * 53 + 11) % 255 - 127
        xs.append(x.astype(np.int8))
        ys.append(y.astype(np.int8))
        exp.append(int((x * y).sum()))
    return dims, xs, ys, exp


def main():
    dims, xs, ys, exp = i8_patterns()
    np.savez(os.path.join(HERE, "i8_dot.npz"), dims=np.array(dims),
             x=np.concatenate(xs), y=np.concatenate(ys), expected=np.array(exp, np.int64))
    ux = [(x.astype(np.int64) + 128).astype(np.uint8) for x in xs]
    uy = [(y.astype(np.int64) + 128).astype(np.uint8) for y in ys]
    uexp = [int((a.astype(np.int64) * b.astype(np.int64)).sum()) for a, b in zip(ux, uy)]
    np.savez(os.path.join(HERE, "u8_dot.npz"), dims=np.array(dims), x=np.concatenate(ux),
             y=np.concatenate(uy), expected=np.array(uexp, np.int64))

    flat = {
        "l2": {"base": [[0.0] * 4, [1.0] * 4, [2.0] * 4], "query": [1.1] * 4, "k": 2,
               "expected_ids": [1, 2], "expected_dist": [0.04, 3.24], "atol": 1e-3},
        "ip": {"base": [[1.0] * 4, [2.0] * 4, [3.0] * 4], "query": [1.0] * 4, "k": 2,
               "expected_ids": [2, 1], "expected_dist": [12.0, 8.0], "atol": 1e-3},
    }
    with open(os.path.join(HERE, "flat_known.json"), "w") as f:
        json.dump(flat, f, indent=1)

    rows = np.array([[((r * 17 + c * 29) % 251) / 251.0 + (2.0 if c == r % 8 else 0.0)
                      for c in range(8)] for r in range(256)], np.float32)
    np.savez(os.path.join(HERE, "ivfpq_rows.npz"), rows=rows)

    rng = np.random.default_rng(7)
    base = rng.integers(0, 4, size=(300, 16)).astype(np.float32)
    base[150:] = base[:150]  # exact duplicates: equal distances at different ids
    q = rng.integers(0, 4, size=(20, 16)).astype(np.float32)
    d2 = ((q[:, None, :].astype(np.float64) - base[None, :, :]) ** 2).sum(-1)
    k = 25
    ids = np.stack([np.lexsort((np.arange(base.shape[0]), row))[:k] for row in d2])
    dist = np.take_along_axis(d2, ids, 1)
    np.savez(os.path.join(HERE, "ties.npz"), base=base, q=q, k=k, ids=ids.astype(np.int64),
             dist=dist.astype(np.float32))


if __name__ == "__main__":
    main()
