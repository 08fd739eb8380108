"""CPU oracle of the learned scheduler's network (test infrastructure only: the tests use it to
check the HIP training and inference kernels; nothing in the product imports it).

A 12 -> 64 -> 64 -> 5 multilayer perceptron with ReLU hidden layers, the parameter layout of
``dyng_hip.h`` (DYNG_HIP_MLP_*): W1[64][12], b1[64], W2[64][64], b2[64], W3[5][64], b3[5], then the
input normalization x_mean[12], x_scale[12] (x_n = (x - mean) * scale). The loss is the squared
error summed over the outputs whose target is known, divided by B * 5 for a minibatch of B: a
NaN target (a schedule not measured on that batch) contributes nothing. The reference project has no such network, so
nothing here restates reference code; the backward pass is pinned by finite differences
(tests/test_oracle.py) and the BF16 forward by rounding each stage as IEEE bfloat16 does (round
to nearest, ties to even, on the upper 16 bits of binary32).
"""

from __future__ import annotations

import numpy as np

IN, HIDDEN, OUT = 12, 64, 5
W1, B1 = 0, HIDDEN * IN
W2 = B1 + HIDDEN
B2 = W2 + HIDDEN * HIDDEN
W3 = B2 + HIDDEN
B3 = W3 + OUT * HIDDEN
TRAINED = B3 + OUT
MEAN = TRAINED
SCALE = MEAN + IN
PARAMS = SCALE + IN


def unpack(p: np.ndarray):
    p = np.asarray(p)
    return (p[W1:B1].reshape(HIDDEN, IN), p[B1:W2], p[W2:B2].reshape(HIDDEN, HIDDEN), p[B2:W3],
            p[W3:B3].reshape(OUT, HIDDEN), p[B3:TRAINED], p[MEAN:SCALE], p[SCALE:PARAMS])


def forward(p: np.ndarray, x: np.ndarray, dtype=np.float64):
    """(y, cache) for inputs x (B x 12)."""
    w1, b1, w2, b2, w3, b3, mean, scale = (a.astype(dtype) for a in unpack(p))
    xn = (np.asarray(x, dtype) - mean) * scale
    z1 = xn @ w1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ w2.T + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ w3.T + b3
    return y, (xn, h1, h2)


def loss(p, x, t, dtype=np.float64) -> float:
    y, _ = forward(p, x, dtype)
    t = np.asarray(t, dtype)
    e = np.where(np.isnan(t), 0.0, y - np.nan_to_num(t))
    return float(np.sum(e ** 2) / e.size)


def gradients(p: np.ndarray, x: np.ndarray, t: np.ndarray, dtype=np.float64) -> np.ndarray:
    """d loss / d (trained parameters), in the parameter layout."""
    w1, b1, w2, b2, w3, b3, mean, scale = (a.astype(dtype) for a in unpack(p))
    y, (xn, h1, h2) = forward(p, x, dtype)
    B = y.shape[0]
    t = np.asarray(t, dtype)
    dy = np.where(np.isnan(t), 0.0, 2.0 * (y - np.nan_to_num(t)) / (B * OUT))
    g = np.zeros(TRAINED, dtype)
    g[W3:B3] = (dy.T @ h2).reshape(-1)
    g[B3:TRAINED] = dy.sum(0)
    dh2 = (dy @ w3) * (h2 > 0)
    g[W2:B2] = (dh2.T @ h1).reshape(-1)
    g[B2:W3] = dh2.sum(0)
    dh1 = (dh2 @ w2) * (h1 > 0)
    g[W1:B1] = (dh1.T @ xn).reshape(-1)
    g[B1:W2] = dh1.sum(0)
    return g


def bf16(a) -> np.ndarray:
    """Round binary32 values to bfloat16 (nearest, ties to even), returned as binary32."""
    f = np.ascontiguousarray(np.asarray(a, np.float32))
    bits = f.view(np.uint32).astype(np.uint64)
    nan = (bits & 0x7F800000) == 0x7F800000
    nan &= (bits & 0x007FFFFF) != 0
    rounded = ((bits + 0x7FFF + ((bits >> 16) & 1)) >> 16) << 16
    out = np.where(nan, (bits | 0x400000) & 0xFFFF0000, rounded).astype(np.uint32)
    return out.view(np.float32)


def forward_bf16(p: np.ndarray, x: np.ndarray) -> np.ndarray:
    """The BF16 inference of the kernels: weights, the normalized input and each hidden layer
    rounded to bfloat16, every dot product accumulated in binary32."""
    w1, b1, w2, b2, w3, b3, mean, scale = unpack(np.asarray(p, np.float32))
    xn = bf16((np.asarray(x, np.float32) - mean) * scale)
    h1 = bf16(np.maximum(xn @ bf16(w1).T + bf16(b1), 0).astype(np.float32))
    h2 = bf16(np.maximum(h1 @ bf16(w2).T + bf16(b2), 0).astype(np.float32))
    return (h2 @ bf16(w3).T + bf16(b3)).astype(np.float32)


def batch_features(out_degree, parents, ins_src, del_src, del_dst) -> dict:
    """The scheduler's raw batch features on G_t (before the batch): the deletions (u, v) that hit
    the old tree (parents[v] == u), the operations whose tail u is a vertex of the graph
    ("touched"), the mean of log2(1 + out-degree(u)) over them (each term rounded to 1/1024, as the
    kernel accumulates it in integers) and the largest such out-degree. `out_degree` is the row
    length of every vertex, `parents` the old tree (or None)."""
    deg = np.asarray(out_degree, np.int64)
    n = len(deg)
    ins_src = np.asarray(ins_src, np.int64)
    del_src = np.asarray(del_src, np.int64)
    del_dst = np.asarray(del_dst, np.int64)
    hits = 0
    if parents is not None and len(del_src):
        par = np.asarray(parents, np.int64)
        ok = (del_dst >= 0) & (del_dst < len(par))
        hits = int(np.count_nonzero(par[del_dst[ok]] == del_src[ok]))
    tails = np.concatenate([del_src, ins_src])
    tails = tails[(tails >= 0) & (tails < n)]
    d = deg[tails]
    terms = np.floor(np.log2(1.0 + d.astype(np.float32)).astype(np.float32) * np.float32(1024.0) + np.float32(0.5))
    touched = len(tails)
    return {"tree_deletions": hits, "touched": touched,
            "log_degree_sum": int(terms.astype(np.int64).sum()),
            "mean_log_degree": float(terms.sum() / 1024.0 / touched) if touched else 0.0,
            "max_degree": int(d.max()) if touched else 0}
