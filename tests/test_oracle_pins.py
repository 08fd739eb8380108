"""Pins of the oracle paths that the golden trees leave open: the engines' distance-only mode
(sequential.cpp:247-268 recover_parents, openmp.cpp's unpack), the 47-bit packing boundary and
the lowest-id canonicalization of imported trees (sssp.cpp:383-403).

None of the reference's fixtures runs the distance-only mode (their weights are at most 100), so
the engines are checked against `tests.cases.brute_tree` (scipy's Dijkstra and the lowest-id
tight in-neighbour, computed with numpy), against the brute-force subtree union for
`invalidated` and against the changed pairs for `affected`. The packing boundary has closed forms
(a path of equal weights). CPU only.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests import cases

INF = oracle.INF
W_TIE = 2_000_000_000


def _csr(n, src, dst, w):
    w = np.asarray(w, np.int64)
    return oracle.build_from_edges(n, src, dst, w.reshape(len(w), -1), oracle.Props.mosp_compatible())


def _edges(g: oracle.Csr, k: int = 0):
    return np.repeat(np.arange(g.n), np.diff(g.row_ptr)), g.col.astype(np.int64), g.w[k].astype(np.int64)


def _check_update(g, t, old_d, old_p, a, c, k=0):
    """The tree after an update against the brute-force tree of the new graph, `invalidated`
    against the subtree union on the old tree, `affected` against the changed pairs."""
    s, d, w = _edges(a.g, k)
    bd, bp = cases.brute_tree(a.g.n, s, d, w, t.source)
    assert np.array_equal(t.dist, bd), "distances"
    assert np.array_equal(t.parent, bp), "parents"
    cf, ct = a.changes(k)
    roots = {int(v) for u, v in zip(cf, ct) if old_p[v] == u}
    assert c.invalidated == cases.subtree_union(old_p, roots)
    assert c.affected == int(((old_d != t.dist) | (old_p != t.parent)).sum())


def test_brute_tree_matches_dijkstra_on_small_graphs():
    # the brute force itself, against the oracle's Dijkstra (pinned by the golden init trees) on
    # graphs with many ties and parallel edges
    rng = np.random.default_rng(11)
    for _ in range(20):
        n = int(rng.integers(2, 40))
        m = int(rng.integers(0, 6 * n))
        s, d, w = rng.integers(0, n, m), rng.integers(0, n, m), rng.integers(1, 4, m)
        g = oracle.build_from_edges(n, s, d, w.reshape(-1, 1), oracle.Props(allow_parallel=True))
        od, op = oracle.dijkstra(g, 0)
        bd, bp = cases.brute_tree(n, *_edges(g), 0)
        assert np.array_equal(od, bd) and np.array_equal(op, bp)


@pytest.mark.parametrize("engine", ["seq", "omp"])
def test_distance_only_canonicalizes_non_canonical_input(engine):
    # sssp_test.cpp:239-273: 272 x 272 one-way grid, every weight 2 * 10^9 (n = 73984 needs 17
    # parent bits; (n - 1) * w does not fit beside them). The imported tree takes the left
    # neighbour where the upper one (lower id) is also tight; deleting (0, 1) must give the
    # canonical tree of the new graph.
    side = 272
    n, s, d, w = cases.grid_one_way(side, W_TIE)
    assert oracle.device_parent_bits(n, W_TIE) == 0 and not oracle.packs_parents(n, W_TIE)
    g = _csr(n, s, d, w)
    t, c0 = oracle.compute(g, 0, engine=engine, threads=4)
    assert not c0.packed_parents
    bd, bp = cases.brute_tree(n, s, d, w, 0)
    assert np.array_equal(t.dist, bd) and np.array_equal(t.parent, bp)
    p = bp.copy().reshape(side, side)
    p[1:, 1:] = np.arange(n, dtype=np.int32).reshape(side, side)[1:, 1:] - 1
    t = oracle.Tree(0, bd.copy(), p.ravel().copy())
    assert not np.array_equal(t.parent, bp)
    old_d, old_p = t.dist.copy(), t.parent.copy()
    g2, a, c = oracle.update(g, t, [], [], [], [0], [1], props=oracle.Props.mosp_compatible(), engine=engine,
                             threads=4)
    assert not c.packed_parents
    _check_update(g2, t, old_d, old_p, a, c)


@pytest.mark.parametrize("engine", ["seq", "omp"])
def test_distance_only_large_weight_ties(engine):
    # sssp_test.cpp:275-315 and sssp_cuda_test.cpp:559-617: 320 x 320 grid, arcs both ways,
    # objective 0 of weight 2 * 10^9 (two tight in-neighbours almost everywhere), objective 1
    # random in [1, 2^31 - 1]; rounds of tree-edge deletions and random insertions.
    side = 320
    n, s, d = cases.grid_two_way(side)
    rng = np.random.default_rng(20260927)
    w = np.stack([np.full(len(s), W_TIE, np.int64), rng.integers(1, 2**31, len(s))], 1)
    g0 = oracle.build_from_edges(n, s, d, w, oracle.Props.mosp_compatible())
    for k in (0, 1):
        g = g0
        t, c0 = oracle.compute(g, 0, k, engine=engine, threads=4)
        assert not c0.packed_parents
        bd, bp = cases.brute_tree(n, *_edges(g, k), 0)
        assert np.array_equal(t.dist, bd) and np.array_equal(t.parent, bp)
        for _ in range(2):
            v = rng.integers(1, n, 500)
            v = v[t.parent[v] >= 0]
            a_, c_ = rng.integers(0, n, 500), rng.integers(0, n, 500)
            keep = a_ != c_
            ins_w = np.stack([np.full(keep.sum(), W_TIE, np.int64), rng.integers(1, 2**31, keep.sum())], 1)
            old_d, old_p = t.dist.copy(), t.parent.copy()
            g, a, c = oracle.update(g, t, a_[keep], c_[keep], ins_w, t.parent[v], v,
                                    props=oracle.Props.mosp_compatible(), engine=engine, threads=4)
            assert not c.packed_parents
            _check_update(g, t, old_d, old_p, a, c, k)


@pytest.mark.parametrize("engine", ["seq", "omp"])
def test_packing_boundary_of_47_distance_bits(engine):
    # sssp_cuda_test.cpp:501-553: n = 2^17 - 1 needs 17 parent bits; on a path of weight
    # W = 2^30 + 2^14 the bound (n - 1) W = 2^47 - 2^15 fits in 47 bits (the device packs), the host
    # rule counts one more edge and keeps distances only. n W does not fit: the back edge n-1 -> 1
    # never offers. Closed forms: d(v) = v W, parent(v) = v - 1.
    n, W = (1 << 17) - 1, (1 << 30) + (1 << 14)
    assert oracle.device_parent_bits(n, W) == 17 and not oracle.packs_parents(n, W)
    vid = np.arange(n, dtype=np.int64)
    for name, back, ins in (("pull", False, (n - 1, 1, W)), ("push", True, (n - 2, n - 1, W - 1)),
                            ("from-scratch", True, None)):
        g = _csr(n, *cases.path(n, W, back))
        t, _ = oracle.compute(g, 0, engine=engine, threads=4)
        assert np.array_equal(t.dist, vid * W), name
        assert np.array_equal(t.parent, (vid - 1).astype(np.int32)), name
        if ins is None:
            continue
        old_d = t.dist.copy()
        _, _, c = oracle.update(g, t, [ins[0]], [ins[1]], [[ins[2]]], [], [],
                                props=oracle.Props.mosp_compatible(), engine=engine, threads=4)
        expect = vid * W
        if name == "push":
            expect[n - 1] -= 1
        assert np.array_equal(t.dist, expect), name
        assert np.array_equal(t.parent, (vid - 1).astype(np.int32)), name
        assert c.affected == int((old_d != t.dist).sum()) == (1 if name == "push" else 0)


def test_canonicalize_gives_the_lowest_id_tight_parents():
    # sssp.cpp:383-403 (MOSP's canonicalizeTree). A valid tree whose parents are random tight
    # in-neighbours comes out as the canonical tree: the oracle's Dijkstra tree (pinned by the
    # golden init trees) and the brute-force lowest-id rule.
    rng = np.random.default_rng(3)
    planted = 0
    for _ in range(30):
        n = int(rng.integers(5, 200))
        m = int(rng.integers(n, 8 * n))
        s, d, w = rng.integers(0, n, m), rng.integers(0, n, m), rng.integers(1, 4, m)
        g = oracle.build_from_edges(n, s, d, w.reshape(-1, 1), oracle.Props(allow_parallel=True))
        dist, par = oracle.dijkstra(g, 0)
        p = cases.random_tight_parents(rng, n, *_edges(g), dist, par, 0)
        planted += int((p != par).sum())
        assert oracle.check_tree(g, 0, dist, p)["ok"]
        out = oracle.canonicalize(g, 0, dist, p)
        assert np.array_equal(out, par)
        assert np.array_equal(out, cases.brute_tree(n, *_edges(g), 0)[1])
        assert np.array_equal(oracle.canonicalize(g, 0, dist, par), par)  # canonical stays canonical
    assert planted > 100  # the planted trees were not canonical already


@pytest.mark.parametrize("engine", ["seq", "omp"])
def test_non_canonical_input_keeps_tie_parents_like_mosp(engine):
    # sssp_test.cpp:317-349: parent[3] = 2 although 1 is also tight; inserting 0 -> 4 (weight 1)
    # offers 3 another tie through 4, which it does not take (4 > 2).
    edges = np.array([(0, 1, 1), (0, 2, 1), (1, 3, 1), (2, 3, 1), (0, 4, 5), (4, 3, 1)], np.int64)
    g = _csr(5, edges[:, 0], edges[:, 1], edges[:, 2])
    t = oracle.Tree(0, np.array([0, 1, 1, 2, 5], np.int64), np.array([-1, 0, 0, 2, 0], np.int32))
    g2, _, c = oracle.update(g, t, [0], [4], [[1]], [], [], props=oracle.Props.mosp_compatible(), engine=engine)
    assert t.dist.tolist() == [0, 1, 1, 2, 1]
    assert t.parent.tolist() == [-1, 0, 0, 2, 0]
    assert c.affected == 1
    assert oracle.check_tree(g2, 0, t.dist, t.parent)["ok"]
    fresh, _ = oracle.compute(g2, 0, engine=engine)
    assert fresh.dist.tolist() == t.dist.tolist() and fresh.parent.tolist() == [-1, 0, 0, 1, 0]
    # canonicalized on import, the update equals compute()
    c_parent = oracle.canonicalize(g, 0, np.array([0, 1, 1, 2, 5]), np.array([-1, 0, 0, 2, 0]))
    assert c_parent.tolist() == [-1, 0, 0, 1, 0]
    t2 = oracle.Tree(0, np.array([0, 1, 1, 2, 5], np.int64), c_parent)
    oracle.update(g, t2, [0], [4], [[1]], [], [], props=oracle.Props.mosp_compatible(), engine=engine)
    assert t2.parent.tolist() == fresh.parent.tolist()
