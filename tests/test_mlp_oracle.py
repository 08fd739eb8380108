"""The CPU oracle of the scheduler's network, pinned by mathematics: finite differences for the
backward pass and the IEEE bfloat16 rounding rule for the BF16 forward pass."""

from __future__ import annotations

import numpy as np

from oracle import mlp


def _params(rng):
    p = rng.normal(0, 0.3, mlp.PARAMS)
    p[mlp.MEAN:mlp.SCALE] = rng.normal(0, 1, mlp.IN)
    p[mlp.SCALE:] = rng.uniform(0.5, 2.0, mlp.IN)
    return p


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(rng)
    x = rng.normal(0, 2, (9, mlp.IN))
    t = rng.normal(0, 1, (9, mlp.OUT))
    g = mlp.gradients(p, x, t)
    h = 1e-6
    for q in rng.choice(mlp.TRAINED, 60, replace=False).tolist() + [mlp.B1, mlp.B2, mlp.B3, mlp.TRAINED - 1]:
        e = np.zeros(mlp.PARAMS)
        e[q] = h
        fd = (mlp.loss(p + e, x, t) - mlp.loss(p - e, x, t)) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 + 1e-5 * abs(fd), (q, fd, g[q])


def test_gradients_with_unknown_targets_match_finite_differences():
    """NaN targets (schedules not measured on a batch) drop out of the loss and its gradient."""
    rng = np.random.default_rng(4)
    p = _params(rng)
    x = rng.normal(0, 2, (11, mlp.IN))
    t = rng.normal(0, 1, (11, mlp.OUT))
    t[rng.random(t.shape) < 0.7] = np.nan
    g = mlp.gradients(p, x, t)
    assert np.isfinite(g).all()
    h = 1e-6
    for q in rng.choice(mlp.TRAINED, 60, replace=False).tolist() + [mlp.B3, mlp.TRAINED - 1]:
        e = np.zeros(mlp.PARAMS)
        e[q] = h
        fd = (mlp.loss(p + e, x, t) - mlp.loss(p - e, x, t)) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 + 1e-5 * abs(fd), (q, fd, g[q])
    # an output with no known target in the minibatch gets no gradient on its own weights
    t2 = t.copy()
    t2[:, 3] = np.nan
    g2 = mlp.gradients(p, x, t2)
    assert np.all(g2[mlp.W3 + 3 * mlp.HIDDEN:mlp.W3 + 4 * mlp.HIDDEN] == 0) and g2[mlp.B3 + 3] == 0


def test_planted_gradient_mistake_is_caught():
    """A backward pass that forgets the ReLU mask disagrees with finite differences."""
    rng = np.random.default_rng(2)
    p = _params(rng)
    x = rng.normal(0, 2, (6, mlp.IN))
    t = rng.normal(0, 1, (6, mlp.OUT))
    w1, b1, w2, b2, w3, b3, mean, scale = mlp.unpack(p)
    y, (xn, h1, h2) = mlp.forward(p, x)
    dy = 2.0 * (y - t) / (6 * mlp.OUT)
    wrong = ((dy @ w3).T @ h1).reshape(-1)  # no mask on dh2
    right = mlp.gradients(p, x, t)[mlp.W2:mlp.B2]
    assert not np.allclose(wrong, right)


def test_bf16_rounding_rule():
    one = np.float32(1.0)
    half_ulp = np.float32(1.0 + 2.0 ** -8)          # exactly between 1 and 1 + 2^-7: ties to even -> 1
    above = np.float32(1.0 + 2.0 ** -8 + 2.0 ** -20)  # just above the tie -> up
    odd_tie = np.float32(1.0 + 2.0 ** -7 + 2.0 ** -8)  # tie with an odd lower neighbour -> up to 1 + 2^-6
    got = mlp.bf16([one, half_ulp, above, odd_tie, -3.0, 0.0])
    assert got.tolist() == [1.0, 1.0, 1.0 + 2.0 ** -7, 1.0 + 2.0 ** -6, -3.0, 0.0]
    assert np.isnan(mlp.bf16([np.nan]))[0]


def test_bf16_forward_is_close_to_fp64():
    rng = np.random.default_rng(3)
    p = _params(rng)
    x = rng.normal(0, 2, (50, mlp.IN))
    y64, _ = mlp.forward(p, x)
    yb = mlp.forward_bf16(p, x)
    assert np.max(np.abs(yb - y64)) <= 0.05 * (1 + np.max(np.abs(y64)))


def test_batch_features_against_a_loop():
    """oracle.mlp.batch_features against a plain loop over the operations (its definition written
    out), and a planted mistake (tree hits tested as parents[u] == v) that the loop catches."""
    import math

    from oracle import mlp as om

    rng = np.random.default_rng(4)
    n = 50
    deg = rng.integers(0, 300, n)
    parents = np.where(rng.random(n) < 0.8, rng.integers(0, n, n), -1)
    for _ in range(20):
        ins_s = rng.integers(0, n + 5, int(rng.integers(0, 40)))
        del_s = rng.integers(0, n, int(rng.integers(0, 40)))
        del_d = rng.integers(0, n + 3, len(del_s))
        hits = sum(1 for u, v in zip(del_s, del_d) if v < n and parents[v] == u)
        tails = [u for u in list(del_s) + list(ins_s) if 0 <= u < n]
        terms = [math.floor(float(np.float32(np.log2(np.float32(1 + deg[u])))) * 1024 + 0.5) for u in tails]
        got = om.batch_features(deg, parents, ins_s, del_s, del_d)
        assert got["tree_deletions"] == hits and got["touched"] == len(tails)
        assert got["log_degree_sum"] == sum(terms)
        assert got["max_degree"] == (max(deg[u] for u in tails) if tails else 0)
        wrong = sum(1 for u, v in zip(del_s, del_d) if u < n and v < n and parents[u] == v)
        if hits != wrong:
            assert got["tree_deletions"] != wrong
