"""The CPU oracle of MOSP's combined solve against the reference's committed combined fixtures
(cpp/tests/data/mosp_combined, copied as data under tests/golden/mosp_combined): for every
preference case, the batch applied, the K canonical trees of the updated graph, their combined
graph, its canonical tree and the path costs (every weight column of the graph, as MOSP's
mospCosts.txt holds them) equal the fixture files; and the thesis worked example."""

from __future__ import annotations

import os

import numpy as np
import pytest

import oracle
from oracle import mosp

from . import fixture_io as fio


def _updated_graph(inp: str):
    prefix = os.path.join(inp, "graphCsr")
    K = fio.num_weight_columns(prefix)
    row, col, w = fio.read_csr(prefix, K)
    g = oracle.Csr(row, col, w)
    b = fio.read_batch(inp, K)
    return oracle.apply_batch(g, b[0], b[1], b[2], b[3], b[4]).g


@pytest.mark.parametrize("name,inp,init,K,prefs", fio.combined_cases(), ids=[c[0] for c in fio.combined_cases()])
def test_combined_solve_matches_the_fixtures(name, inp, init, K, prefs):
    g = _updated_graph(inp)
    trees = [oracle.dijkstra(g, 0, k)[1] for k in range(K)]
    d, p, _ = mosp.solve(trees, 0, prefs)
    want = os.path.join(fio.GOLDEN, "mosp_combined", name, "combined")
    assert np.array_equal(d, fio.read_distances(os.path.join(want, "distancesCsr.txt"), g.n)), name
    assert np.array_equal(p, fio.read_parents(os.path.join(want, "SSSPTreeCsr.txt"), g.n)), name
    costs = fio.read_costs(os.path.join(want, "mospCosts.txt"), g.n)
    assert np.array_equal(mosp.path_costs(g, p, 0, costs.shape[1]), costs), name


def test_thesis_worked_example():
    """Thesis Chapter 4 (python/dyng/mosp.py docstring): with Pref (4, 1, 4) the MOSP path to
    vertex 6 costs (15, 3, 20) after the batch."""
    src = [0, 0, 1, 2, 2, 3, 4, 4, 4, 5]
    dst = [1, 2, 3, 1, 3, 4, 1, 5, 6, 6]
    w = np.array([[2, 1, 5], [4, 1, 1], [2, 4, 2], [10, 15, 2], [5, 16, 3], [1, 1, 1], [4, 3, 2],
                  [1, 2, 2], [5, 6, 2], [1, 1, 1]])
    g = oracle.build_from_edges(7, np.array(src), np.array(dst), w)
    g = oracle.apply_batch(g, np.array([3, 1]), np.array([5, 5]), np.array([[10, 2, 12], [12, 1, 14]]),
                           np.array([1, 4]), np.array([3, 1])).g
    trees = [oracle.dijkstra(g, 0, k)[1] for k in range(3)]
    _, p, _ = mosp.solve(trees, 0, [4, 1, 4])
    assert mosp.path_costs(g, p, 0, 3)[6].tolist() == [15, 3, 20]


def test_combine_counts_an_edge_once_and_weighs_it_by_its_trees():
    """Closed form: two objectives that share the tree edge (0, 1) give it base - L/p0 - L/p1
    once; an edge of one tree weighs base - its term."""
    parents = [np.array([-1, 0, 0]), np.array([-1, 0, 1])]
    g, wsum, base = mosp.combine(parents, 0, [1, 2])
    # L = 2, base = 2 * 3 = 6, terms = (2, 1)
    assert base == 6
    edges = {(int(u), int(v)): int(x) for u in range(3) for v, x in
             zip(g.col[g.row_ptr[u]:g.row_ptr[u + 1]], g.w[0][g.row_ptr[u]:g.row_ptr[u + 1]])}
    assert edges == {(0, 1): 3, (0, 2): 4, (1, 2): 5}
    assert wsum == 12


def test_planted_combine_mistake_is_caught():
    """Counting a shared tree edge once per tree (a dropped dedup) changes the fixture's tree."""
    name, inp, init, K, prefs = fio.combined_cases()[0]
    g = _updated_graph(inp)
    trees = [oracle.dijkstra(g, 0, k)[1] for k in range(K)]
    cg, _, _ = mosp.combine(trees, 0, prefs)
    # the planted variant: every objective adds its own copy of the edge with only its own term
    terms, base, _ = mosp.terms_and_base(prefs, K)
    src, dst, w = [], [], []
    for v in range(g.n):
        for k in range(K):
            if v != 0 and trees[k][v] >= 0:
                src.append(int(trees[k][v]))
                dst.append(v)
                w.append(base - terms[k])
    bad = oracle.build_from_edges(g.n, np.array(src), np.array(dst), np.array(w).reshape(-1, 1),
                                  oracle.Props(allow_parallel=True))
    d_bad, _ = oracle.dijkstra(bad, 0, 0)
    d_good, _ = oracle.dijkstra(cg, 0, 0)
    assert not np.array_equal(d_bad, d_good)


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_fast_combine_and_costs_equal_the_loops(seed):
    """The whole-array combine and path costs (the CPU baseline's) equal the loop versions on
    random multigraphs (parallel edges, unreachable vertices) and on every fixture case."""
    rng = np.random.default_rng(seed)
    n, m, K = 400, 1600, 3
    src, dst = rng.integers(0, n, m), rng.integers(0, n, m)
    w = rng.integers(1, 20, (m, K))
    g = oracle.build_from_edges(n, src, dst, w, oracle.Props.mosp_compatible())
    trees = [oracle.dijkstra(g, 0, k)[1] for k in range(K)]
    prefs = [[], [1, 2, 3], [4, 1, 4], [7, 7, 2]][seed - 1]
    a, sa, ba = mosp.combine(trees, 0, prefs)
    b, sb, bb = mosp.combine_fast(trees, 0, prefs)
    assert (sa, ba) == (sb, bb)
    assert np.array_equal(a.row_ptr, b.row_ptr) and np.array_equal(a.col, b.col) and np.array_equal(a.w, b.w)
    d, p = oracle.dijkstra(a, 0, 0)
    assert np.array_equal(mosp.path_costs(g, p, 0, K), mosp.path_costs_fast(g, p, 0, K))


def test_fast_costs_on_the_fixtures_and_a_missing_edge():
    for name, inp, init, K, prefs in fio.combined_cases():
        g = _updated_graph(inp)
        trees = [oracle.dijkstra(g, 0, k)[1] for k in range(K)]
        _, p, _ = mosp.solve(trees, 0, prefs)
        want = fio.read_costs(os.path.join(fio.GOLDEN, "mosp_combined", name, "combined", "mospCosts.txt"), g.n)
        assert np.array_equal(mosp.path_costs_fast(g, p, 0, want.shape[1]), want), name
    g = oracle.build_from_edges(3, np.array([0]), np.array([1]), np.array([[1]]))
    with pytest.raises(ValueError):
        mosp.path_costs_fast(g, np.array([-1, 0, 1]), 0, 1)


def test_update_many_equals_one_update_per_tree():
    """oracle.update_many (dyng::update_each): one apply, K repairs = K separate updates."""
    rng = np.random.default_rng(9)
    n, m, K = 500, 4000, 3
    src, dst = rng.integers(0, n, m), rng.integers(0, n, m)
    w = rng.integers(1, 50, (m, K))
    g = oracle.build_from_edges(n, src, dst, w)
    many = [oracle.compute(g, s, k)[0] for s, k in ((0, 0), (3, 1), (0, 2))]
    alone = [oracle.compute(g, s, k)[0] for s, k in ((0, 0), (3, 1), (0, 2))]
    ins_s, ins_d = rng.integers(0, n + 2, 200), rng.integers(0, n + 2, 200)
    ins_w = rng.integers(1, 50, (200, K))
    rows = np.repeat(np.arange(g.n), np.diff(g.row_ptr))
    pick = rng.integers(0, g.m, 150)
    g2, _, cs = oracle.update_many(g, many, ins_s, ins_d, ins_w, rows[pick], g.col[pick])
    for t, tm, c in zip(alone, many, cs):
        _, _, ca = oracle.update(g, t, ins_s, ins_d, ins_w, rows[pick], g.col[pick])
        assert np.array_equal(t.dist, tm.dist) and np.array_equal(t.parent, tm.parent)
        assert (c.invalidated, c.affected) == (ca.invalidated, ca.affected)
        d, p = oracle.dijkstra(g2, t.source, t.objective)
        assert np.array_equal(tm.dist, d) and np.array_equal(tm.parent, p)
