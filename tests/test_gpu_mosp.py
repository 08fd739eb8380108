"""Parity of the HIP path's MOSP combined solve (dyng_amd.mosp over dyng_hip_mosp_combined) with
the reference's committed combined fixtures (cpp/tests/data/mosp_combined, copied as data under
tests/golden/mosp_combined) and with the oracle (oracle/mosp.py). Needs an MI355X: every test here
is marked gpu.

The bar is bit-exactness: the combined distances, the MOSP tree, the path costs, the combined edge
count and `affected` (pairs (distance, parent) changed by an update) equal the fixtures' and the
oracle's.
"""

from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np
import pytest

import oracle
from oracle import mosp as omosp
from tests import fixture_io as fio

pytestmark = pytest.mark.gpu

INF = oracle.INF


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dg():
    import dyng_amd

    dyng_amd._lib.lib()  # loud failure if the HIP library is missing
    assert dyng_amd.device_count() > 0, "no HIP device visible"
    return dyng_amd


def _all_column_costs(dg, g, r) -> np.ndarray:
    """Path costs over every weight column of g (MOSP's mospCosts.txt), through the C entry point
    with r's K trees."""
    from dyng_amd import _lib

    n, cols = g.num_vertices, g.num_weights
    K = r.num_objectives
    ptrs = (C.c_void_p * K)(*[r.tree(k).device_arrays()[1] for k in range(K)])
    prefs = r.options.preferences
    pref_arr = (C.c_int32 * K)(*prefs) if prefs else None
    dist = dg.DeviceArray(np.full(n, INF, np.int64))
    par = dg.DeviceArray(np.full(n, -1, np.int32))
    costs = dg.DeviceArray(np.zeros((n, cols), np.int64))
    st = _lib.MospStats()
    _lib.check(_lib.lib().dyng_hip_mosp_combined(g._handle, r.source, K, ptrs, pref_arr, 0, dist.ptr, par.ptr,
                                                 cols, costs.ptr, None, C.byref(st)))
    assert np.array_equal(dist.to_numpy(), r.combined_distances)
    assert np.array_equal(par.to_numpy(), r.combined_parents)
    return costs.to_numpy()


def _check_against_fixture(dg, g, r, name):
    want = os.path.join(fio.GOLDEN, "mosp_combined", name, "combined")
    n = g.num_vertices
    d, p = r.combined_distances, r.combined_parents
    assert fio.distances_text(d) == fio.read_text(os.path.join(want, "distancesCsr.txt")), name
    assert fio.parents_text(p) == fio.read_text(os.path.join(want, "SSSPTreeCsr.txt")), name
    costs = fio.read_costs(os.path.join(want, "mospCosts.txt"), n)
    K = r.num_objectives
    assert np.array_equal(r.path_costs, costs[:, :K]), name
    assert np.array_equal(_all_column_costs(dg, g, r), costs), name


@pytest.mark.parametrize("name,inp,init,K,prefs", fio.combined_cases(), ids=[c[0] for c in fio.combined_cases()])
def test_golden_combined_update_from_the_original_trees(dg, name, inp, init, K, prefs):
    """`dyng mosp update --init` (python/tests/test_cli.py:689-721): the original trees imported
    without canonicalization, the batch applied once, the combined files equal the fixture's."""
    _say(f"[mosp golden] {name} K={K} prefs={prefs}")
    prefix = os.path.join(inp, "graphCsr")
    cols = fio.num_weight_columns(prefix)
    row, col, w = fio.read_csr(prefix, cols)
    g = dg.Graph.from_csr(row, col, w.T, properties=dg.GraphProperties.mosp_compatible())
    n = g.num_vertices
    ds = [fio.read_distances(os.path.join(init, f"obj{k}", "distancesOriginal.txt"), n) for k in range(K)]
    ps = [fio.read_parents(os.path.join(init, f"obj{k}", "SSSPTreeOriginal.txt"), n) for k in range(K)]
    r = dg.mosp.Result.from_arrays(g, 0, ds, ps, canonicalize=False, preferences=prefs, num_objectives=K)
    ins_src, ins_dst, ins_w, del_src, del_dst = fio.read_batch(inp, cols)
    st = dg.mosp.update(g, dg.EdgeBatch(insert=(ins_src, ins_dst, ins_w), delete=(del_src, del_dst)), r)
    assert r.graph_version == g.version == 1
    assert len(st.objectives) == K and st.converged and not st.fallback_used
    _check_against_fixture(dg, g, r, name)


@pytest.mark.parametrize("name,inp,init,K,prefs", fio.combined_cases(), ids=[c[0] for c in fio.combined_cases()])
def test_golden_combined_compute_then_update(dg, name, inp, init, K, prefs):
    """mosp.compute on the original graph, then mosp.update: the same combined files; the update's
    counters equal the oracle's (affected against the compute's combined tree)."""
    _say(f"[mosp compute+update] {name}")
    prefix = os.path.join(inp, "graphCsr")
    cols = fio.num_weight_columns(prefix)
    row, col, w = fio.read_csr(prefix, cols)
    g = dg.Graph.from_csr(row, col, w.T, properties=dg.GraphProperties.mosp_compatible())
    r = dg.mosp.compute(g, 0, preferences=prefs, num_objectives=K)
    og = oracle.Csr(row, col, w)
    t0 = [oracle.dijkstra(og, 0, k)[1] for k in range(K)]
    d0, p0, m0 = omosp.solve(t0, 0, prefs)
    assert np.array_equal(r.combined_distances, d0) and np.array_equal(r.combined_parents, p0)
    assert r.combined_edges == m0
    ins_src, ins_dst, ins_w, del_src, del_dst = fio.read_batch(inp, cols)
    st = dg.mosp.update(g, dg.EdgeBatch(insert=(ins_src, ins_dst, ins_w), delete=(del_src, del_dst)), r)
    _check_against_fixture(dg, g, r, name)
    og1 = oracle.apply_batch(og, ins_src, ins_dst, ins_w, del_src, del_dst, oracle.Props.mosp_compatible()).g
    t1 = [oracle.dijkstra(og1, 0, k)[1] for k in range(K)]
    d1, p1, m1 = omosp.solve(t1, 0, prefs)
    assert st.combined_edges == m1
    assert st.affected == int(np.count_nonzero((d1 != d0) | (p1 != p0)))
    assert st.preference_scale == r.preference_scale == omosp.preference_scale(prefs, K)


def test_thesis_worked_example(dg):
    """Thesis Chapter 4 (python/dyng/mosp.py:22-37): with Pref (4, 1, 4) the MOSP path to vertex 6
    costs (15, 3, 20) after the batch."""
    src = [0, 0, 1, 2, 2, 3, 4, 4, 4, 5]
    dst = [1, 2, 3, 1, 3, 4, 1, 5, 6, 6]
    w = [[2, 1, 5], [4, 1, 1], [2, 4, 2], [10, 15, 2], [5, 16, 3], [1, 1, 1], [4, 3, 2],
         [1, 2, 2], [5, 6, 2], [1, 1, 1]]
    g = dg.Graph.from_edges(src, dst, np.array(w), properties=dg.GraphProperties.mosp_compatible())
    paths = dg.mosp.compute(g, 0, preferences=[4, 1, 4])
    batch = dg.EdgeBatch(insert=([3, 1], [5, 5], [[10, 2, 12], [12, 1, 14]]), delete=([1, 4], [3, 1]))
    dg.mosp.update(g, batch, paths)
    assert paths.path_costs[6].tolist() == [15, 3, 20]


def _random_graph(rng, n, m, K, wmax):
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    w = rng.integers(1, wmax + 1, (m, K))
    return src, dst, w


@pytest.mark.parametrize("seed,n,m,K,prefs,simple", [
    (1, 300, 1500, 3, [4, 1, 4], False),
    (2, 1000, 6000, 2, [3, 5], False),
    (3, 2000, 9000, 4, [], False),
    (4, 500, 1500, 3, [1, 2, 7], False),
    (5, 1500, 9000, 3, [4, 1, 4], True),
    (6, 800, 3000, 5, [1, 2, 3, 4, 5], True),
    # more trees than the combined diff holds in registers (its generic path)
    (7, 400, 2400, 10, [1, 2, 3, 4, 5, 6, 7, 8, 9, 10], False),
])
def test_random_updates_match_the_oracle(dg, seed, n, m, K, prefs, simple):
    """Random multigraphs (parallel edges, self-loops, unreachable parts) or simple graphs with
    sorted rows (whose tree edges the costs look up in the in-rows), several batches with
    insertions, deletions, weight changes and vertex growth: after every update the combined
    arrays, the costs and the counters equal the oracle's."""
    _say(f"[mosp random] seed={seed} n={n} m={m} K={K} prefs={prefs} simple={simple}")
    rng = np.random.default_rng(seed)
    src, dst, w = _random_graph(rng, n, m, K, 20)
    op = oracle.Props() if simple else oracle.Props.mosp_compatible()
    gp = dg.GraphProperties() if simple else dg.GraphProperties.mosp_compatible()
    og = oracle.build_from_edges(n, src, dst, w, op)
    g = dg.Graph.from_edges(src, dst, w, num_vertices=n, properties=gp)
    r = dg.mosp.compute(g, 0, preferences=prefs)
    prev_d, prev_p, _ = omosp.solve([oracle.dijkstra(og, 0, k)[1] for k in range(K)], 0, prefs)
    assert np.array_equal(r.combined_distances, prev_d) and np.array_equal(r.combined_parents, prev_p)
    for step in range(4):
        nv = og.n
        grow = 3 if step == 2 else 0
        ni, nd = int(rng.integers(1, 60)), int(rng.integers(1, 60))
        ins_s = rng.integers(0, nv + grow, ni)
        ins_d = rng.integers(0, nv + grow, ni)
        ins_w = rng.integers(1, 21, (ni, K))
        # deletions of existing edges (and a few absent pairs)
        e = rng.integers(0, max(og.m, 1), nd)
        rows = np.repeat(np.arange(og.n), np.diff(og.row_ptr))
        del_s = np.concatenate([rows[e], rng.integers(0, nv, 3)])
        del_d = np.concatenate([og.col[e], rng.integers(0, nv, 3)])
        st = dg.mosp.update(g, dg.EdgeBatch(insert=(ins_s, ins_d, ins_w), delete=(del_s, del_d)), r)
        og = oracle.apply_batch(og, ins_s, ins_d, ins_w, del_s, del_d, op).g
        trees = [oracle.dijkstra(og, 0, k)[1] for k in range(K)]
        d, p, mc = omosp.solve(trees, 0, prefs)
        for k in range(K):
            assert np.array_equal(r.parents(k), trees[k]), (step, k)
        assert np.array_equal(r.combined_distances, d), step
        assert np.array_equal(r.combined_parents, p), step
        assert np.array_equal(r.path_costs, omosp.path_costs(og, p, 0, K)), step
        assert st.combined_edges == mc
        pd = np.full(og.n, INF, np.int64)
        pp = np.full(og.n, -1, np.int64)
        pd[:len(prev_d)] = prev_d
        pp[:len(prev_p)] = prev_p
        assert st.affected == int(np.count_nonzero((d != pd) | (p != pp))), step
        prev_d, prev_p = d, p


def test_options_and_errors(dg):
    g = dg.Graph.from_edges([0, 1], [1, 2], np.array([[1, 2], [3, 4]]),
                            properties=dg.GraphProperties.mosp_compatible())
    with pytest.raises(dg.InvalidArgumentError):
        dg.mosp.compute(g, 0, preferences=[1, 0])
    with pytest.raises(dg.InvalidArgumentError):
        dg.mosp.compute(g, 0, preferences=[1, 2, 3])
    with pytest.raises(dg.InvalidArgumentError):
        dg.mosp.compute(g, 0, preferences=[1 << 20, 3])
    with pytest.raises(dg.InvalidArgumentError):
        dg.mosp.compute(g, 5)
    r = dg.mosp.compute(g, 0, compute_path_costs=False)
    with pytest.raises(dg.InvalidArgumentError):
        _ = r.path_costs
    r.set_options(compute_path_costs=True)
    dg.mosp.update(g, dg.EdgeBatch(insert=([0], [2], [[1, 1]])), r)
    assert r.path_costs.tolist() == [[0, 0], [1, 2], [1, 1]]
    with pytest.raises(dg.InvalidArgumentError):
        r.set_options(preferences=[2, 1])
    c = r.clone()
    g2 = dg.Graph.from_edges([0], [1], np.array([[1, 1]]), properties=dg.GraphProperties.mosp_compatible())
    with pytest.raises(dg.StaleResultError):
        dg.mosp.update(g2, dg.EdgeBatch(insert=([1], [0], [[1, 1]])), c)
    assert np.array_equal(c.combined_parents, r.combined_parents)


# ------------------------------------------------------------------------------------------------
# dyng_amd.update: one batch, several results (cpp/include/dyng/update.hpp:263-364)
# ------------------------------------------------------------------------------------------------


def test_update_several_results_equals_each_computed_alone(dg):
    """Trees of different objectives and sources plus a mosp result, one batch applied once: each
    equals what compute gives on the new graph, and the graph's version grows by one."""
    _say("[update_each] mixed results")
    rng = np.random.default_rng(77)
    n, m, K = 3000, 20000, 3
    src, dst, w = _random_graph(rng, n, m, K, 30)
    op = oracle.Props()
    og = oracle.build_from_edges(n, src, dst, w, op)
    g = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    t0 = dg.sssp.compute(g, 0, objective=0)
    t1 = dg.sssp.compute(g, 17, objective=2)
    t2 = dg.sssp.compute(g, 5, objective=1, delta=7)
    mr = dg.mosp.compute(g, 0, preferences=[2, 3, 1])
    for b in range(3):
        ni, nd = 400, 300
        ins_s, ins_d = rng.integers(0, n, ni), rng.integers(0, n, ni)
        ins_w = rng.integers(1, 31, (ni, K))
        rows = np.repeat(np.arange(og.n), np.diff(og.row_ptr))
        pick = rng.integers(0, og.m, nd)
        del_s, del_d = rows[pick], og.col[pick]
        version = g.version
        stats = dg.update(g, dg.EdgeBatch(insert=(ins_s, ins_d, ins_w), delete=(del_s, del_d)), t0, t1, mr, t2)
        assert g.version == version + 1
        assert len(stats) == 4 and isinstance(stats[2], dg.mosp.Stats) and isinstance(stats[0], dg.sssp.Stats)
        og = oracle.apply_batch(og, ins_s, ins_d, ins_w, del_s, del_d, op).g
        for t, (s, k) in zip((t0, t1, t2), ((0, 0), (17, 2), (5, 1))):
            d, p = oracle.dijkstra(og, s, k)
            assert np.array_equal(t.distances, d) and np.array_equal(t.parents, p), (b, s, k)
            assert t.graph_version == g.version
        trees = [oracle.dijkstra(og, 0, k)[1] for k in range(K)]
        d, p, mc = omosp.solve(trees, 0, [2, 3, 1])
        assert np.array_equal(mr.combined_distances, d) and np.array_equal(mr.combined_parents, p), b
        assert np.array_equal(mr.path_costs, omosp.path_costs(og, p, 0, K)), b
        assert stats[2].combined_edges == mc
        for s in stats[:2] + stats[3:]:
            assert s.batch.inserted_edges + s.batch.updated_edges == stats[2].batch.inserted_edges + \
                stats[2].batch.updated_edges


def test_update_checks_everything_before_the_batch(dg):
    """No result, a result twice, a stale result, an insertion weight below 1: an error, and the
    graph and every result unchanged."""
    g = dg.Graph.from_edges([0, 1, 2], [1, 2, 0], np.array([[1, 2], [1, 2], [1, 2]]))
    a = dg.sssp.compute(g, 0)
    b = dg.sssp.compute(g, 1, objective=1)
    batch = dg.EdgeBatch(delete=([2], [0]))
    with pytest.raises(dg.InvalidArgumentError):
        dg.update(g, batch)
    with pytest.raises(dg.InvalidArgumentError):
        dg.update(g, batch, a, b, a)
    stale = dg.sssp.compute(g, 0)
    dg.sssp.update(g, dg.EdgeBatch(insert=([0], [2], [[5, 5]])), stale)  # g moves on; a and b go stale
    assert g.version == 1
    with pytest.raises(dg.StaleResultError):
        dg.update(g, batch, stale, a)
    assert g.version == 1 and g.num_edges == 4
    a = dg.sssp.compute(g, 0)
    b = dg.sssp.compute(g, 1, objective=1)
    with pytest.raises(dg.InvalidArgumentError):
        dg.update(g, dg.EdgeBatch(insert=([1], [0], [[3, 0]])), a, b)  # objective 1 weight 0
    assert g.version == 1 and g.num_edges == 4
    st = dg.update(g, batch, a, b)
    assert g.version == 2 and g.num_edges == 3 and len(st) == 2
    assert a.distances.tolist() == [0, 1, 2] and b.distances.tolist()[1] == 0


def test_update_many_results_reuse_workspaces(dg):
    """Nine results (more than the workspace slots: runs of one slot follow one another, two are
    in flight at once) through three batches: each tree equals Dijkstra's on the new graph."""
    _say("[update_each] nine results")
    rng = np.random.default_rng(91)
    n, m, K = 2500, 15000, 3
    src, dst, w = _random_graph(rng, n, m, K, 25)
    op = oracle.Props()
    og = oracle.build_from_edges(n, src, dst, w, op)
    g = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    picks = [(int(rng.integers(0, n)), int(rng.integers(0, K))) for _ in range(9)]
    trees = [dg.sssp.compute(g, s, objective=k) for s, k in picks]
    for b in range(3):
        ni, nd = 300, 300
        ins_s, ins_d = rng.integers(0, n, ni), rng.integers(0, n, ni)
        ins_w = rng.integers(1, 26, (ni, K))
        rows = np.repeat(np.arange(og.n), np.diff(og.row_ptr))
        pick = rng.integers(0, og.m, nd)
        del_s, del_d = rows[pick], og.col[pick]
        stats = dg.update(g, dg.EdgeBatch(insert=(ins_s, ins_d, ins_w), delete=(del_s, del_d)), *trees)
        og = oracle.apply_batch(og, ins_s, ins_d, ins_w, del_s, del_d, op).g
        for t, (s, k), st in zip(trees, picks, stats):
            d, p = oracle.dijkstra(og, s, k)
            assert np.array_equal(t.distances, d) and np.array_equal(t.parents, p), (b, s, k)
            assert t.graph_version == g.version and st.converged


def test_path_costs_down_a_deep_chain(dg):
    """The path doubling of the costs on a tree deeper than its grid covers in one stride: a chain
    of 600,000 vertices (the costs' grid is capped at 2048 x 256 = 524,288 threads, so threads take
    two vertices each and the doubling runs 20 rounds through the settled states), with a few
    shortcuts that the trees do not take, then a batch that cuts the chain and adds a bypass. The
    combined tree and every objective's cost equal the oracle's after each step."""
    _say("[mosp deep chain] n=600000 K=2")
    n, K = 600_000, 2
    rng = np.random.default_rng(11)
    src = np.arange(n - 1, dtype=np.int64)
    dst = src + 1
    w = rng.integers(1, 10, (n - 1, K))
    # shortcuts far heavier than the chain below them: never on a shortest path
    sc_u = rng.integers(0, n - 1000, 50)
    sc_v = sc_u + rng.integers(2, 1000, 50)
    src = np.concatenate([src, sc_u])
    dst = np.concatenate([dst, sc_v])
    w = np.concatenate([w, np.full((50, K), 20_000, np.int64)])
    props = dg.GraphProperties.mosp_compatible()
    g = dg.Graph.from_edges(src, dst, w, num_vertices=n, properties=props)
    og = oracle.build_from_edges(n, src, dst, w, oracle.Props.mosp_compatible())
    r = dg.mosp.compute(g, 0, preferences=[1, 2])

    def check(og):
        trees = [oracle.dijkstra(og, 0, k)[1] for k in range(K)]
        d, p, _ = omosp.solve(trees, 0, [1, 2])
        assert np.array_equal(r.combined_distances, d) and np.array_equal(r.combined_parents, p)
        assert np.array_equal(r.path_costs, omosp.path_costs_fast(og, p, 0, K))

    check(og)
    # cut the chain at 300,000 and bypass it from 299,990 to 300,005
    ins_s, ins_d, ins_w = [299_990], [300_005], [[7, 7]]
    del_s, del_d = [300_000], [300_001]
    dg.mosp.update(g, dg.EdgeBatch(insert=(ins_s, ins_d, ins_w), delete=(del_s, del_d)), r)
    og = oracle.apply_batch(og, ins_s, ins_d, ins_w, del_s, del_d, oracle.Props.mosp_compatible()).g
    check(og)
