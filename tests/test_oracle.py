"""The CPU oracle pinned against the reference's golden data, hand cases and closed forms.

These tests need no GPU. They pin oracle/ before it is trusted as the checker of the HIP path:
  * MOSP applyChangeBatch goldens (cpp/tests/data/mosp_graph_io/<case>/applied): the applied CSR,
    the transposed CSR and the weight-increase flags, byte for byte;
  * MOSP-OpenMP golden trees (cpp/tests/data/mosp_sssp, 34 cases x objectives): compute() equals
    the Dijkstra init tree, update() from the original's files equals the updated tree, and
    `invalidated` equals stats.txt -- for the sequential and the OpenMP engine (1, 2, 4 threads),
    as cpp/tests/algorithms/sssp/sssp_fixture_test.cpp:108-190 runs them;
  * hand cases of cpp/tests/algorithms/sssp/sssp_test.cpp and python/tests/test_sssp.py;
  * closed forms: Dijkstra against brute-force Bellman-Ford, update == compute from canonical
    trees, `invalidated` == the brute-force subtree union, `affected` == the changed pairs;
  * planted mistakes (in the oracle's inputs, on the CPU) that check_tree must catch.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

import oracle
from tests import fixture_io as fio

INF = oracle.INF


def _k_of(values_path: str) -> int:
    with open(values_path) as f:
        for line in f:
            if line.split():
                return len(line.split())
    return 1


# ------------------------------------------------------------------------------------------------
# MOSP apply goldens
# ------------------------------------------------------------------------------------------------


@pytest.mark.parametrize("case", fio.apply_cases())
def test_apply_matches_mosp_apply_goldens(case):
    base = os.path.join(fio.GOLDEN, "mosp_graph_io", case)
    K = _k_of(os.path.join(base, "graphCsrValues.txt"))
    row, col, w = fio.read_csr(os.path.join(base, "graphCsr"), K)
    g = oracle.Csr(row, col, w)
    ins_src, ins_dst, ins_w, del_src, del_dst = fio.read_batch(base, K)
    a = oracle.apply_batch(g, ins_src, ins_dst, ins_w, del_src, del_dst,
                           oracle.Props.mosp_compatible())
    erow, ecol, ew = fio.read_csr(os.path.join(base, "applied", "graphCsr"), K)
    assert np.array_equal(a.g.row_ptr, erow)
    assert np.array_equal(a.g.col, ecol)
    assert np.array_equal(a.g.w, ew)
    t = oracle.transpose(a.g)
    trow, tcol, tw = fio.read_csr(os.path.join(base, "applied", "graphCsrTransposed"), K)
    assert np.array_equal(t.row_ptr, trow) and np.array_equal(t.col, tcol)
    assert np.array_equal(t.w, tw)
    inc = fio._ints(os.path.join(base, "applied", "graphCsrWeightIncrease.txt"))
    expect = np.array(inc, np.uint8).reshape(len(inc), K) if inc else np.zeros((0, K), np.uint8)
    assert np.array_equal(a.weight_increased, expect)


# ------------------------------------------------------------------------------------------------
# MOSP-OpenMP golden trees
# ------------------------------------------------------------------------------------------------

ENGINES = [("seq", 1), ("omp", 1), ("omp", 2), ("omp", 4)]


@pytest.mark.parametrize("name,inp,K", fio.sssp_cases(), ids=[c[0] for c in fio.sssp_cases()])
def test_engines_match_mosp_sssp_goldens(name, inp, K):
    row, col, w = fio.read_csr(os.path.join(inp, "graphCsr"), K)
    g = oracle.Csr(row, col, w)
    batch = fio.read_batch(inp, K)
    expected = os.path.join(fio.GOLDEN, "mosp_sssp", name)
    props = oracle.Props.mosp_compatible()
    for engine, threads in ENGINES:
        for k in range(K):
            obj = f"obj{k}"
            tree, _ = oracle.compute(g, 0, k, engine, threads=threads)
            assert fio.distances_text(tree.dist) == fio.read_text(
                os.path.join(expected, "init", obj, "distancesOriginal.txt")), (engine, threads, k)
            assert fio.parents_text(tree.parent) == fio.read_text(
                os.path.join(expected, "init", obj, "SSSPTreeOriginal.txt")), (engine, threads, k)
            # update() from the original's initial files (canonicalize = false) == mosp
            d0 = fio.read_distances(os.path.join(expected, "init", obj, "distancesOriginal.txt"), g.n)
            p0 = fio.read_parents(os.path.join(expected, "init", obj, "SSSPTreeOriginal.txt"), g.n)
            t = oracle.Tree(0, d0, p0, k)
            _, _, c = oracle.update(g, t, *batch, props=props, engine=engine, threads=threads)
            assert fio.distances_text(t.dist) == fio.read_text(
                os.path.join(expected, "updated", obj, "distancesUpdated.txt")), (engine, threads, k)
            assert fio.parents_text(t.parent) == fio.read_text(
                os.path.join(expected, "updated", obj, "SSSPTreeUpdated.txt")), (engine, threads, k)
            assert c.invalidated == fio.expected_invalidated(name, k), (engine, threads, k)


# ------------------------------------------------------------------------------------------------
# Hand cases (cpp/tests/algorithms/sssp/sssp_test.cpp, python/tests/test_sssp.py)
# ------------------------------------------------------------------------------------------------


def _graph(n, edges, props=None):
    e = np.array(edges, np.int64).reshape(-1, 3)
    return oracle.build_from_edges(n, e[:, 0], e[:, 1], e[:, 2:3],
                                   props or oracle.Props.mosp_compatible())


@pytest.mark.parametrize("engine", ["seq", "omp"])
def test_ties_case(engine):
    g = _graph(5, [(0, 1, 5), (0, 2, 1), (2, 3, 1), (1, 4, 2), (3, 4, 2)])
    t, _ = oracle.compute(g, 0, engine=engine)
    assert t.dist.tolist() == [0, 5, 1, 2, 4] and t.parent.tolist() == [-1, 0, 0, 2, 3]
    g2, a, c = oracle.update(g, t, [2], [1], [[1]], [], [], engine=engine)
    assert t.dist.tolist() == [0, 2, 1, 2, 4] and t.parent.tolist() == [-1, 2, 0, 2, 1]
    assert c.invalidated == 0 and c.affected == 2 and a.summary["inserted_edges"] == 1
    assert c.packed_parents


@pytest.mark.parametrize("engine", ["seq", "omp"])
def test_disconnect_does_not_count_to_infinity(engine):
    g = _graph(4, [(0, 1, 1), (0, 3, 50), (1, 2, 1), (2, 1, 1), (3, 1, 50)])
    t, _ = oracle.compute(g, 0, engine=engine)
    _, _, c = oracle.update(g, t, [], [], [], [0], [1], engine=engine)
    assert t.dist.tolist() == [0, 100, 101, 50] and t.parent.tolist() == [-1, 3, 1, 0]
    assert c.invalidated == 2 and c.affected == 2


@pytest.mark.parametrize("engine", ["seq", "omp"])
def test_delete_all_then_reinsert(engine):
    g = _graph(5, [(0, 1, 3), (1, 2, 4), (0, 3, 2), (3, 4, 7)])
    t, _ = oracle.compute(g, 0, engine=engine)
    g, _, c = oracle.update(g, t, [], [], [], [0, 1, 0, 3], [1, 2, 3, 4], engine=engine)
    assert g.m == 0
    assert t.dist.tolist() == [0, INF, INF, INF, INF] and t.parent.tolist() == [-1] * 5
    assert c.invalidated == 4 and c.affected == 4
    g, _, _ = oracle.update(g, t, [0, 4], [4, 2], [[9], [1]], [], [], engine=engine)
    assert t.dist.tolist() == [0, INF, 10, INF, 9] and t.parent.tolist() == [-1, -1, 4, -1, 0]


@pytest.mark.parametrize("engine", ["seq", "omp"])
def test_empty_batch(engine):
    g = _graph(4, [(0, 1, 2), (1, 2, 2), (0, 2, 4), (2, 3, 1)])
    t, _ = oracle.compute(g, 0, engine=engine)
    d, p = t.dist.copy(), t.parent.copy()
    _, _, c = oracle.update(g, t, [], [], [], [], [], engine=engine)
    assert c.affected == 0 and c.invalidated == 0 and c.iterations == 0
    assert np.array_equal(t.dist, d) and np.array_equal(t.parent, p)


@pytest.mark.parametrize("engine", ["seq", "omp"])
def test_weight_increase_on_tree_edge(engine):
    g = _graph(3, [(0, 1, 1), (1, 2, 1), (0, 2, 5)])
    t, _ = oracle.compute(g, 0, engine=engine)
    _, a, c = oracle.update(g, t, [0], [1], [[10]], [], [], engine=engine)
    assert c.invalidated == 2 and a.summary["updated_edges"] == 1
    assert t.dist.tolist() == [0, 10, 5] and t.parent.tolist() == [-1, 0, 0]


@pytest.mark.parametrize("engine", ["seq", "omp"])
def test_parallel_edges_and_self_loops(engine):
    g = _graph(3, [(0, 1, 7), (0, 1, 3), (1, 1, 1), (1, 2, 1)])
    t, _ = oracle.compute(g, 0, engine=engine)
    assert t.dist.tolist() == [0, 3, 4]
    g, _, c = oracle.update(g, t, [], [], [], [0], [1], engine=engine)
    assert c.invalidated == 2 and t.dist.tolist() == [0, 3, 4]
    g, _, _ = oracle.update(g, t, [], [], [], [0], [1], engine=engine)
    assert t.dist.tolist() == [0, INF, INF]


@pytest.mark.parametrize("engine", ["seq", "omp"])
def test_vertex_growth(engine):
    g = _graph(3, [(0, 1, 2), (1, 2, 2)], oracle.Props())
    t, _ = oracle.compute(g, 0, engine=engine)
    g, a, _ = oracle.update(g, t, [2, 5], [5, 3], [[1], [4]], [], [], props=oracle.Props(),
                            engine=engine)
    assert g.n == 6 and a.summary["num_vertices_after"] == 6
    assert t.dist.tolist() == [0, 2, 4, 9, INF, 5] and t.parent.tolist() == [-1, 0, 1, 5, -1, 2]
    assert oracle.check_tree(g, 0, t.dist, t.parent)["ok"]


@pytest.mark.parametrize("engine", ["seq", "omp"])
def test_undirected_graph(engine):
    props = oracle.Props(directed=False)
    g = _graph(4, [(0, 1, 1), (1, 2, 1), (2, 3, 1)], props)
    t, _ = oracle.compute(g, 3, engine=engine)
    assert t.dist.tolist() == [3, 2, 1, 0]
    oracle.update(g, t, [0], [3], [[5]], [1], [2], props=props, engine=engine)
    assert t.dist.tolist() == [5, 6, 1, 0] and t.parent.tolist() == [3, 0, 3, -1]


@pytest.mark.parametrize("engine", ["seq", "omp"])
def test_multi_objective_uses_the_chosen_column(engine):
    e = np.array([(0, 1, 1, 10), (1, 2, 1, 10), (0, 2, 5, 1)])
    g = oracle.build_from_edges(3, e[:, 0], e[:, 1], e[:, 2:4], oracle.Props.mosp_compatible())
    t1, _ = oracle.compute(g, 0, 1, engine)
    t0, _ = oracle.compute(g, 0, 0, engine)
    assert t0.dist.tolist() == [0, 1, 2] and t1.dist.tolist() == [0, 10, 1]
    g1, _, c0 = oracle.update(g, t0, [0], [2], [[5, 30]], [], [], engine=engine)
    _, _, c1 = oracle.update(g, t1, [0], [2], [[5, 30]], [], [], engine=engine)
    assert c0.invalidated == 0 and c1.invalidated == 1
    assert t1.dist.tolist() == [0, 10, 20] and t1.parent.tolist() == [-1, 0, 1]


@pytest.mark.parametrize("engine", ["seq", "omp"])
def test_python_update_matches_compute(engine):
    g = _graph(4, [(0, 1, 4), (0, 2, 1), (1, 2, 1), (2, 3, 5)], oracle.Props())
    t, _ = oracle.compute(g, 0, engine=engine)
    assert t.dist.tolist() == [0, 4, 1, 6] and t.parent.tolist() == [-1, 0, 0, 2]
    g2, a, c = oracle.update(g, t, [1, 3], [3, 4], [[1], [2]], [0], [2], props=oracle.Props(),
                             engine=engine)
    assert a.summary["inserted_edges"] == 2 and a.summary["deleted_edges"] == 1
    assert c.invalidated == 2
    fresh, _ = oracle.compute(g2, 0, engine=engine)
    assert t.dist.tolist() == fresh.dist.tolist() == [0, 4, 5, 5, 7]
    assert t.parent.tolist() == fresh.parent.tolist()


def test_distance_only_mode_closed_form():
    # n = 73984 > 2^16 needs 17 parent bits; with weight 2^31 - 1 the bound does not fit beside
    # them (sssp_test.cpp:239-316 uses the same shape).
    n, w = 73984, 2**31 - 1
    assert oracle.device_parent_bits(n, w) == 0 and not oracle.packs_parents(n, w)
    assert oracle.device_parent_bits(n, 100) == 17 and oracle.packs_parents(n, 100)
    # the packing window (ADR 0029): the CUDA rule packs, the OpenMP rule does not
    bits = 17
    max_distance = ((1 << 64) - 1 >> bits) - 1
    wmax = max_distance // (n - 1)
    assert oracle.device_parent_bits(n, wmax) == 17
    assert not oracle.packs_parents(n, wmax)


def test_default_delta_closed_form():
    assert oracle.default_delta(1_048_576, 65_536, 1_048_576 * 50) == int(32 * 50 / 16)
    assert oracle.default_delta(0, 10, 0) == 1
    assert oracle.default_delta(10, 10, 10) == 32


# ------------------------------------------------------------------------------------------------
# Closed forms on random graphs
# ------------------------------------------------------------------------------------------------


def _bellman_ford(g: oracle.Csr, source: int, k: int = 0):
    n = g.n
    dist = [INF] * n
    dist[source] = 0
    src = np.repeat(np.arange(n), np.diff(g.row_ptr))
    for _ in range(n):
        changed = False
        for e in range(g.m):
            u, v, w = int(src[e]), int(g.col[e]), int(g.w[k][e])
            if dist[u] < INF and dist[u] + w < dist[v]:
                dist[v] = dist[u] + w
                changed = True
        if not changed:
            break
    parent = [-1] * n
    for v in range(n):
        if v == source or dist[v] >= INF:
            continue
        best = -1
        for e in range(g.m):
            u = int(src[e])
            if int(g.col[e]) == v and dist[u] < INF and dist[u] + int(g.w[k][e]) == dist[v]:
                best = u if best < 0 else min(best, u)
        parent[v] = best
    return dist, parent


def _random_graph(rng, n, m, wmax, props=None):
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    w = rng.integers(1, wmax + 1, m)
    return oracle.build_from_edges(n, src, dst, w.reshape(-1, 1), props or oracle.Props())


@pytest.mark.parametrize("seed", range(6))
def test_dijkstra_matches_bellman_ford(seed):
    rng = np.random.default_rng(seed)
    g = _random_graph(rng, int(rng.integers(2, 25)), int(rng.integers(0, 70)), int(rng.integers(1, 6)))
    d, p = oracle.dijkstra(g, 0)
    bd, bp = _bellman_ford(g, 0)
    assert d.tolist() == bd and p.tolist() == bp


def _subtree_union(parent, roots):
    n = len(parent)
    children = [[] for _ in range(n)]
    for v, p in enumerate(parent):
        if p >= 0:
            children[p].append(v)
    seen = set()
    stack = list(roots)
    while stack:
        x = stack.pop()
        if x in seen:
            continue
        seen.add(x)
        stack.extend(children[x])
    return seen


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("engine", ["seq", "omp"])
def test_random_update_equals_compute_and_brute_force_counts(seed, engine):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(5, 60))
    g = _random_graph(rng, n, int(rng.integers(n, 6 * n)), 9)
    t, _ = oracle.compute(g, 0, engine=engine)
    for _ in range(3):
        ni, nd = int(rng.integers(0, 12)), int(rng.integers(0, 12))
        ins = (rng.integers(0, n, ni), rng.integers(0, n, ni), rng.integers(1, 10, (ni, 1)))
        rows = np.repeat(np.arange(g.n), np.diff(g.row_ptr))
        pick = rng.integers(0, max(g.m, 1), nd) if g.m else np.zeros(0, np.int64)
        dels = (rows[pick] if g.m else [], g.col[pick] if g.m else [])
        old_d, old_p = t.dist.copy(), t.parent.copy()
        g2, a, c = oracle.update(g, t, *ins, *dels, engine=engine)
        # invalidated: the union of the subtrees below changed tree edges, on the old tree
        cf, ct = a.changes(0)
        roots = {int(v) for u, v in zip(cf, ct) if old_p[v] == u}
        assert c.invalidated == len(_subtree_union(old_p, roots))
        fresh, _ = oracle.compute(g2, 0, engine=engine)
        assert np.array_equal(t.dist, fresh.dist) and np.array_equal(t.parent, fresh.parent)
        grow = len(t.dist) - len(old_d)
        od = np.concatenate([old_d, np.full(grow, INF, np.int64)])
        op = np.concatenate([old_p, np.full(grow, -1, np.int32)])
        assert c.affected == int(((od != t.dist) | (op != t.parent)).sum())
        assert oracle.check_tree(g2, 0, t.dist, t.parent, require_canonical=True)["ok"]
        g = g2


def test_check_tree_catches_planted_mistakes():
    rng = np.random.default_rng(5)
    g = _random_graph(rng, 30, 150, 5)
    d, p = oracle.dijkstra(g, 0)
    assert oracle.check_tree(g, 0, d, p, require_canonical=True)["ok"]
    reach = [v for v in range(1, 30) if d[v] < INF and p[v] >= 0]
    v = reach[0]
    bad = d.copy()
    bad[v] += 1
    assert oracle.check_tree(g, 0, bad, p)["distance_mismatches"] >= 1
    badp = p.copy()
    badp[v] = (p[v] + 1) % 30
    r = oracle.check_tree(g, 0, d, badp, require_canonical=True)
    assert r["inconsistent_parents"] + r["non_canonical_parents"] >= 1 and not r["ok"]


def test_omp_engine_rejects_parent_cycle():
    g = _graph(4, [(0, 1, 4), (0, 2, 1), (1, 2, 1), (2, 3, 5)], oracle.Props())
    t = oracle.Tree(0, np.array([0, 4, 1, 6], np.int64), np.array([-1, 3, 0, 1], np.int32))
    for engine in ("seq", "omp"):
        tt = oracle.Tree(0, t.dist.copy(), t.parent.copy())
        with pytest.raises(oracle.OracleError, match="cycle"):
            oracle.update(g, tt, [], [], [], [0], [2], engine=engine)
