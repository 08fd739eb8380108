"""Parity of the HIP path (libdyng_hip.so through dyng_amd) with the oracle and the reference's
golden data. Needs an MI355X: every test here is marked gpu.

The bar is bit-exactness: distances (int64), parents, `invalidated`, `affected` and the apply
summary equal the oracle's (the sequential restatement of the reference engine) after every batch;
the device rows equal the oracle's MOSP apply row for row (as multisets ordered by neighbour id).
"""

from __future__ import annotations

import os
import sys

import numpy as np
import pytest

import oracle
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


def _props_pair(dg, directed=True, sorted_rows=True, allow_parallel=False, self_loop="keep",
                growth=True, existing="upsert", deletions_first=True):
    loops = {"keep": 0, "drop": 1, "error": 2}
    ins = {"upsert": 0, "ignore": 2}
    op = oracle.Props(directed=directed, sorted_rows=sorted_rows, allow_parallel=allow_parallel,
                      on_existing_insert=ins[existing], on_self_loop=loops[self_loop],
                      deletions_first=deletions_first, allow_vertex_growth=growth)
    sem = dg.BatchSemantics(on_existing_insert=existing, on_self_loop=self_loop,
                            deletions_first=deletions_first, allow_vertex_growth=growth)
    gp = dg.GraphProperties(directed=directed, order="sorted" if sorted_rows else "append",
                            parallel_edges="allow" if allow_parallel else "forbid", semantics=sem)
    return op, gp


def _sorted_rows(row, col, w, by_weight=False):
    """Order every row by neighbour id, stable (or by (neighbour, weights) for in-rows)."""
    n = len(row) - 1
    rows = np.repeat(np.arange(n), np.diff(row))
    keys = [np.arange(len(col))]
    if by_weight:
        keys = [w[k] for k in range(w.shape[0] - 1, -1, -1)]
    order = np.lexsort(tuple(keys) + (col, rows))
    return row, col[order], w[:, order]


def _assert_same_rows(hg, og: oracle.Csr):
    hr, hc, hw = hg.to_csr()
    orow, ocol, ow = _sorted_rows(og.row_ptr, og.col, og.w)
    assert np.array_equal(hr, orow), "row offsets"
    assert np.array_equal(hc, ocol), "neighbours"
    assert np.array_equal(hw, ow), "weights"
    ir, ic, iw = hg.in_csr()
    t = oracle.transpose(og)
    trow, tcol, tw = _sorted_rows(t.row_ptr, t.col, t.w, by_weight=True)
    _, ic2, iw2 = _sorted_rows(ir, ic, iw, by_weight=True)
    assert np.array_equal(ir, trow) and np.array_equal(ic2, tcol) and np.array_equal(iw2, tw), "in-rows"


# ------------------------------------------------------------------------------------------------
# Golden data of the reference (cpp/tests/data/mosp_sssp, 34 cases)
# ------------------------------------------------------------------------------------------------


@pytest.mark.parametrize("name,inp,K", fio.sssp_cases(), ids=[c[0] for c in fio.sssp_cases()])
def test_golden_compute_and_update(dg, name, inp, K):
    _say(f"[golden] {name} K={K}")
    row, col, w = fio.read_csr(os.path.join(inp, "graphCsr"), K)
    ins_src, ins_dst, ins_w, del_src, del_dst = fio.read_batch(inp, K)
    expected = os.path.join(fio.GOLDEN, "mosp_sssp", name)
    props = dg.GraphProperties.mosp_compatible()
    for k in range(K):
        obj = f"obj{k}"
        g = dg.Graph.from_csr(row, col, w.T, properties=props)
        tree = dg.sssp.compute(g, 0, objective=k)
        d, p = tree.to_numpy()
        assert fio.distances_text(d) == fio.read_text(os.path.join(expected, "init", obj, "distancesOriginal.txt"))
        assert fio.parents_text(p) == fio.read_text(os.path.join(expected, "init", obj, "SSSPTreeOriginal.txt"))
        d0 = fio.read_distances(os.path.join(expected, "init", obj, "distancesOriginal.txt"), g.num_vertices)
        p0 = fio.read_parents(os.path.join(expected, "init", obj, "SSSPTreeOriginal.txt"), g.num_vertices)
        r = dg.sssp.Result.from_arrays(g, 0, d0, p0, canonicalize=False, objective=k)
        batch = dg.EdgeBatch(insert=(ins_src, ins_dst, ins_w), delete=(del_src, del_dst))
        st = dg.sssp.update(g, batch, r)
        d1, p1 = r.to_numpy()
        assert fio.distances_text(d1) == fio.read_text(os.path.join(expected, "updated", obj, "distancesUpdated.txt"))
        assert fio.parents_text(p1) == fio.read_text(os.path.join(expected, "updated", obj, "SSSPTreeUpdated.txt"))
        assert st.invalidated == fio.expected_invalidated(name, k)
        assert r.graph_version == g.version == 1


@pytest.mark.parametrize("case", fio.apply_cases())
def test_golden_device_apply(dg, case):
    _say(f"[apply] {case}")
    base = os.path.join(fio.GOLDEN, "mosp_graph_io", case)
    with open(os.path.join(base, "graphCsrValues.txt")) as f:
        K = next(len(line.split()) for line in f if line.split())
    row, col, w = fio.read_csr(os.path.join(base, "graphCsr"), K)
    g = dg.Graph.from_csr(row, col, w.T, properties=dg.GraphProperties.mosp_compatible())
    ins_src, ins_dst, ins_w, del_src, del_dst = fio.read_batch(base, K)
    summary = g.apply(dg.EdgeBatch(insert=(ins_src, ins_dst, ins_w), delete=(del_src, del_dst)))
    erow, ecol, ew = fio.read_csr(os.path.join(base, "applied", "graphCsr"), K)
    _assert_same_rows(g, oracle.Csr(erow, ecol, ew))
    a = oracle.apply_batch(oracle.Csr(row, col, w), ins_src, ins_dst, ins_w, del_src, del_dst,
                           oracle.Props.mosp_compatible())
    for f in ("inserted_edges", "updated_edges", "deleted_edges", "ignored_deletions"):
        assert getattr(summary, f) == a.summary[f], f


# ------------------------------------------------------------------------------------------------
# Hand cases (cpp/tests/algorithms/sssp/sssp_test.cpp, python/tests/test_sssp.py)
# ------------------------------------------------------------------------------------------------


def _g(dg, n, edges, props=None):
    e = np.array(edges, np.int64).reshape(-1, 3)
    return dg.Graph.from_edges(e[:, 0], e[:, 1], e[:, 2], num_vertices=n,
                               properties=props or dg.GraphProperties.mosp_compatible())


def test_hand_ties(dg):
    g = _g(dg, 5, [(0, 1, 5), (0, 2, 1), (2, 3, 1), (1, 4, 2), (3, 4, 2)])
    r = dg.sssp.compute(g, 0)
    assert r.distances.tolist() == [0, 5, 1, 2, 4] and r.parents.tolist() == [-1, 0, 0, 2, 3]
    st = dg.sssp.update(g, dg.EdgeBatch(insert=([2], [1], [1])), r)
    assert r.distances.tolist() == [0, 2, 1, 2, 4] and r.parents.tolist() == [-1, 2, 0, 2, 1]
    assert st.invalidated == 0 and st.affected == 2 and st.batch.inserted_edges == 1 and st.packed_parents


def test_hand_disconnect(dg):
    g = _g(dg, 4, [(0, 1, 1), (0, 3, 50), (1, 2, 1), (2, 1, 1), (3, 1, 50)])
    r = dg.sssp.compute(g, 0)
    st = dg.sssp.update(g, dg.EdgeBatch(delete=([0], [1])), r)
    assert r.distances.tolist() == [0, 100, 101, 50] and r.parents.tolist() == [-1, 3, 1, 0]
    assert st.invalidated == 2 and st.affected == 2


def test_hand_delete_all_and_reinsert(dg):
    g = _g(dg, 5, [(0, 1, 3), (1, 2, 4), (0, 3, 2), (3, 4, 7)])
    r = dg.sssp.compute(g, 0)
    st = dg.sssp.update(g, dg.EdgeBatch(delete=([0, 1, 0, 3], [1, 2, 3, 4])), r)
    assert g.num_edges == 0
    assert r.distances.tolist() == [0, INF, INF, INF, INF] and r.parents.tolist() == [-1] * 5
    assert st.invalidated == 4 and st.affected == 4
    dg.sssp.update(g, dg.EdgeBatch(insert=([0, 4], [4, 2], [9, 1])), r)
    assert r.distances.tolist() == [0, INF, 10, INF, 9] and r.parents.tolist() == [-1, -1, 4, -1, 0]


def test_hand_empty_batch(dg):
    g = _g(dg, 4, [(0, 1, 2), (1, 2, 2), (0, 2, 4), (2, 3, 1)])
    r = dg.sssp.compute(g, 0)
    d, p = r.to_numpy()
    st = dg.sssp.update(g, dg.EdgeBatch(), r)
    assert st.affected == 0 and st.invalidated == 0 and st.iterations == 0
    d2, p2 = r.to_numpy()
    assert np.array_equal(d, d2) and np.array_equal(p, p2)
    assert g.version == 1 and r.graph_version == 1


def test_hand_weight_increase(dg):
    g = _g(dg, 3, [(0, 1, 1), (1, 2, 1), (0, 2, 5)])
    r = dg.sssp.compute(g, 0)
    st = dg.sssp.update(g, dg.EdgeBatch(insert=([0], [1], [10])), r)
    assert st.invalidated == 2 and st.batch.updated_edges == 1
    assert r.distances.tolist() == [0, 10, 5] and r.parents.tolist() == [-1, 0, 0]


def test_hand_parallel_edges_and_self_loops(dg):
    g = _g(dg, 3, [(0, 1, 7), (0, 1, 3), (1, 1, 1), (1, 2, 1)])
    r = dg.sssp.compute(g, 0)
    assert r.distances.tolist() == [0, 3, 4]
    st = dg.sssp.update(g, dg.EdgeBatch(delete=([0], [1])), r)
    assert st.invalidated == 2 and r.distances.tolist() == [0, 3, 4]
    dg.sssp.update(g, dg.EdgeBatch(delete=([0], [1])), r)
    assert r.distances.tolist() == [0, INF, INF]


def test_hand_vertex_growth(dg):
    g = _g(dg, 3, [(0, 1, 2), (1, 2, 2)], dg.GraphProperties())
    r = dg.sssp.compute(g, 0)
    st = dg.sssp.update(g, dg.EdgeBatch(insert=([2, 5], [5, 3], [1, 4])), r)
    assert g.num_vertices == 6 and st.batch.num_vertices_after == 6
    assert r.distances.tolist() == [0, 2, 4, 9, INF, 5] and r.parents.tolist() == [-1, 0, 1, 5, -1, 2]


def test_hand_undirected(dg):
    g = _g(dg, 4, [(0, 1, 1), (1, 2, 1), (2, 3, 1)], dg.GraphProperties(directed=False))
    r = dg.sssp.compute(g, 3)
    assert r.distances.tolist() == [3, 2, 1, 0]
    dg.sssp.update(g, dg.EdgeBatch(insert=([0], [3], [5]), delete=([1], [2])), r)
    assert r.distances.tolist() == [5, 6, 1, 0] and r.parents.tolist() == [3, 0, 3, -1]


def test_hand_python_quickstart(dg):
    g = dg.Graph.from_edges([0, 0, 1, 2], [1, 2, 2, 3], [4, 1, 1, 5])
    r = dg.sssp.compute(g, 0)
    assert r.distances.tolist() == [0, 4, 1, 6] and r.parents.tolist() == [-1, 0, 0, 2]
    st = dg.sssp.update(g, dg.EdgeBatch(insert=([1, 3], [3, 4], [1, 2]), delete=([0], [2])), r)
    assert st.batch.inserted_edges == 2 and st.batch.deleted_edges == 1 and st.invalidated == 2
    fresh = dg.sssp.compute(g, 0)
    assert r.distances.tolist() == fresh.distances.tolist() == [0, 4, 5, 5, 7]
    assert r.parents.tolist() == fresh.parents.tolist()


def test_errors_and_stale_results(dg):
    g = _g(dg, 4, [(0, 1, 4), (0, 2, 1), (1, 2, 1), (2, 3, 5)], dg.GraphProperties())
    with pytest.raises(dg.InvalidArgumentError, match="out of range"):
        dg.sssp.compute(g, 9)
    r = dg.sssp.compute(g, 0)
    other = dg.sssp.compute(g, 0)
    dg.sssp.update(g, dg.EdgeBatch(delete=([0], [2])), r)
    with pytest.raises(dg.StaleResultError):
        dg.sssp.update(g, dg.EdgeBatch(delete=([0], [1])), other)
    with pytest.raises(dg.InvalidArgumentError):
        dg.sssp.update(g, dg.EdgeBatch(insert=([0], [3], [0])), r)  # weight below 1: nothing changes
    assert g.version == 1
    with pytest.raises(dg.InvalidArgumentError):
        dg.sssp.Result.from_arrays(g, 0, [3, 4, 1, 6], [-1, 0, 0, 2])  # source not at 0
    with pytest.raises(dg.InvalidArgumentError):  # a parent cycle
        dg.sssp.Result.from_arrays(g, 0, [0, 4, 1, 6], [-1, 3, 0, 1], canonicalize=False)
    # an unvalidated cyclic tree: the device check reports the cycle and poisons the result
    bad = dg.sssp.Result.from_arrays(g, 0, [0, 4, INF, 6], [-1, 3, -1, 1], canonicalize=False,
                                     validate_inputs=False)
    with pytest.raises(dg.InvalidArgumentError, match="cycle"):
        dg.sssp.update(g, dg.EdgeBatch(delete=([0], [1])), bad)
    with pytest.raises(dg.StaleResultError):
        bad.to_numpy()


def test_options_and_clone(dg):
    e = np.array([(0, 1, 1, 10), (1, 2, 1, 10), (0, 2, 5, 1)])
    g = dg.Graph.from_edges(e[:, 0], e[:, 1], e[:, 2:4], properties=dg.GraphProperties.mosp_compatible())
    r1 = dg.sssp.compute(g, 0, objective=1, delta=5)
    r0 = dg.sssp.compute(g, 0)
    assert r0.distances.tolist() == [0, 1, 2] and r1.distances.tolist() == [0, 10, 1]
    assert r1.options.objective == 1 and r1.options.delta == 5
    r1.set_options(delta=7)
    assert r1.options.delta == 7
    with pytest.raises(dg.InvalidArgumentError):
        r1.set_options(objective=0)
    c = r0.clone()
    st0 = dg.sssp.update(g, dg.EdgeBatch(insert=([0], [2], [[5, 30]])), r0)
    assert st0.invalidated == 0
    assert c.distances.tolist() == [0, 1, 2] and c.graph_version == 0


# ------------------------------------------------------------------------------------------------
# Random multi-batch streams against the oracle
# ------------------------------------------------------------------------------------------------

STREAM_CASES = [
    # (seed, n, m, wmax, props)
    (1, 40, 160, 9, {}),
    (2, 200, 1200, 100, {}),
    (3, 120, 900, 5, {"allow_parallel": True, "sorted_rows": False}),
    (4, 150, 700, 20, {"directed": False}),
    (5, 300, 2000, 1000, {"self_loop": "drop"}),
    (6, 100, 500, 7, {"existing": "ignore"}),
    (7, 80, 400, 12, {"deletions_first": False, "allow_parallel": True, "sorted_rows": False}),
    (8, 2000, 16000, 255, {}),
    # long rows (past the 16-entry bisection) with several parallel copies of each edge
    (9, 64, 20000, 50, {"allow_parallel": True}),
    (10, 400, 40000, 100, {"allow_parallel": True, "deletions_first": False}),
    # simple graphs where a batch inserts and then deletes a pair, or ignores re-insertions (the
    # engine offers each insertion's head the pair's weights after the batch, as the apply left them)
    (11, 300, 2000, 50, {"deletions_first": False}),
    (12, 200, 1500, 30, {"existing": "ignore", "deletions_first": False}),
    (13, 250, 1800, 40, {"directed": False, "deletions_first": False}),
]


@pytest.mark.parametrize("seed,n,m,wmax,kw", STREAM_CASES, ids=[f"s{c[0]}" for c in STREAM_CASES])
def test_random_streams_match_the_oracle(dg, seed, n, m, wmax, kw):
    _say(f"[stream] seed={seed} n={n} m={m} {kw}")
    rng = np.random.default_rng(seed)
    op, gp = _props_pair(dg, **kw)
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    w = rng.integers(1, wmax + 1, m)
    og = oracle.build_from_edges(n, src, dst, w.reshape(-1, 1), op)
    hg = dg.Graph.from_edges(src, dst, w, num_vertices=n, properties=gp)
    _assert_same_rows(hg, og)
    source = int(rng.integers(0, n))
    ot, _ = oracle.compute(og, source)
    hr = dg.sssp.compute(hg, source)
    d, p = hr.to_numpy()
    assert np.array_equal(d, ot.dist) and np.array_equal(p, ot.parent)
    for b in range(4):
        ni, nd = int(rng.integers(0, m // 8 + 2)), int(rng.integers(0, m // 8 + 2))
        grow = 3 if b == 2 else 0
        ins = (rng.integers(0, n + grow, ni), rng.integers(0, n + grow, ni), rng.integers(1, wmax + 1, ni))
        rows = np.repeat(np.arange(og.n), np.diff(og.row_ptr))
        pick = rng.integers(0, og.m, nd) if og.m else np.zeros(0, np.int64)
        # plus possibly missing pairs, and pairs this batch also inserts (inserted and deleted in
        # one batch, in either order)
        dels = (np.concatenate([rows[pick], rng.integers(0, n, 2), ins[0][:3]]),
                np.concatenate([og.col[pick], rng.integers(0, n, 2), ins[1][:3]]))
        og, a, c = oracle.update(og, ot, ins[0], ins[1], ins[2].reshape(-1, 1), dels[0], dels[1], props=op)
        st = dg.sssp.update(hg, dg.EdgeBatch(insert=ins, delete=dels), hr)
        d, p = hr.to_numpy()
        assert np.array_equal(d, ot.dist), f"distances, batch {b}"
        assert np.array_equal(p, ot.parent), f"parents, batch {b}"
        assert st.invalidated == c.invalidated and st.affected == c.affected
        for f in ("inserted_edges", "updated_edges", "deleted_edges", "ignored_deletions",
                  "dropped_self_loops", "ignored_insertions", "num_vertices_after"):
            assert getattr(st.batch, f) == a.summary[f], f
        _assert_same_rows(hg, og)


@pytest.mark.parametrize("light,chunk,bpc,gb", [(1, 16, 1, 0), (4, 64, 2, 0), (64, 1024, 0, 0), (4096, 256, 0, 1),
                                                (32, 64, 0, 7)])
def test_schedule_knobs_never_change_the_result(dg, light, chunk, bpc, gb):
    _say(f"[tuning] light={light} chunk={chunk} bpc={bpc} grid={gb}")
    rng = np.random.default_rng(99)
    n, m = 3000, 40000
    src = np.concatenate([rng.integers(0, n, m), np.zeros(3000, np.int64)])  # a hub row
    dst = np.concatenate([rng.integers(0, n, m), np.arange(3000)])
    w = rng.integers(1, 50, len(src))
    og = oracle.build_from_edges(n, src, dst, w.reshape(-1, 1))
    hg = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    ot, _ = oracle.compute(og, 0)
    hr = dg.sssp.compute(hg, 0)
    hr.set_tuning(dg.sssp.Tuning(light_degree=light, chunk_edges=chunk, blocks_per_cu=bpc, grid_blocks=gb))
    rows = np.repeat(np.arange(og.n), np.diff(og.row_ptr))
    for _ in range(3):
        pick = rng.integers(0, og.m, 300)
        ins = (rng.integers(0, n, 300), rng.integers(0, n, 300), rng.integers(1, 50, 300))
        dels = (rows[pick], og.col[pick])
        og, a, c = oracle.update(og, ot, ins[0], ins[1], ins[2].reshape(-1, 1), dels[0], dels[1])
        rows = np.repeat(np.arange(og.n), np.diff(og.row_ptr))
        st = dg.sssp.update(hg, dg.EdgeBatch(insert=ins, delete=dels), hr)
        d, p = hr.to_numpy()
        assert np.array_equal(d, ot.dist) and np.array_equal(p, ot.parent)
        assert st.invalidated == c.invalidated and st.affected == c.affected


def test_device_resident_batch(dg):
    """A batch in device memory is applied in place and gives the host batch's result."""
    rng = np.random.default_rng(17)
    n, m = 500, 4000
    src, dst, w = rng.integers(0, n, m), rng.integers(0, n, m), rng.integers(1, 30, m)
    og = oracle.build_from_edges(n, src, dst, w.reshape(-1, 1))
    hg = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    ot, _ = oracle.compute(og, 0)
    hr = dg.sssp.compute(hg, 0)
    rows = np.repeat(np.arange(og.n), np.diff(og.row_ptr))
    pick = rng.integers(0, og.m, 200)
    ins = (rng.integers(0, n, 150), rng.integers(0, n, 150), rng.integers(1, 30, 150))
    dels = (rows[pick], og.col[pick])
    og, a, c = oracle.update(og, ot, ins[0], ins[1], ins[2].reshape(-1, 1), dels[0], dels[1])

    def D(x):
        return dg.DeviceArray(np.asarray(x, np.int32))

    batch = dg.EdgeBatch(insert=(D(ins[0]), D(ins[1]), D(ins[2])), delete=(D(dels[0]), D(dels[1])))
    assert batch.on_device
    st = dg.sssp.update(hg, batch, hr)
    d, p = hr.to_numpy()
    assert np.array_equal(d, ot.dist) and np.array_equal(p, ot.parent)
    assert st.invalidated == c.invalidated and st.affected == c.affected


# ------------------------------------------------------------------------------------------------
# Subtree invalidation: the engine's own tree (top-down walk) and an imported one (pointer jumping)
# ------------------------------------------------------------------------------------------------

def _stream(dg, og, hg, ot, hr, rng, batches, n_ops, wlo, whi, extra_dels=()):
    rows = np.repeat(np.arange(og.n), np.diff(og.row_ptr))
    for b in range(batches):
        pick = rng.integers(0, og.m, n_ops)
        ins = (rng.integers(0, og.n, n_ops), rng.integers(0, og.n, n_ops), rng.integers(wlo, whi, n_ops))
        ds = np.concatenate([rows[pick], np.asarray([e[0] for e in extra_dels], np.int64)])
        dd = np.concatenate([og.col[pick], np.asarray([e[1] for e in extra_dels], np.int64)])
        og, a, c = oracle.update(og, ot, ins[0], ins[1], ins[2].reshape(-1, 1), ds, dd)
        rows = np.repeat(np.arange(og.n), np.diff(og.row_ptr))
        st = dg.sssp.update(hg, dg.EdgeBatch(insert=ins, delete=(ds, dd)), hr)
        d, p = hr.to_numpy()
        assert np.array_equal(d, ot.dist), f"distances, batch {b}"
        assert np.array_equal(p, ot.parent), f"parents, batch {b}"
        assert st.invalidated == c.invalidated and st.affected == c.affected, f"counters, batch {b}"
        extra_dels = ()
    return og


def test_deep_subtrees_are_walked_level_by_level(dg):
    """A chain of unit edges under long shortcuts: deleting chain edges near the source
    invalidates subtrees thousands of levels deep."""
    _say("[deep] chain n=6000")
    rng = np.random.default_rng(5)
    n = 6000
    src = np.concatenate([np.arange(n - 1), rng.integers(0, n, 12000)])
    dst = np.concatenate([np.arange(1, n), rng.integers(0, n, 12000)])
    w = np.concatenate([np.ones(n - 1, np.int64), rng.integers(1500, 4000, 12000)])
    og = oracle.build_from_edges(n, src, dst, w.reshape(-1, 1))
    hg = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    ot, _ = oracle.compute(og, 0)
    hr = dg.sssp.compute(hg, 0)
    _stream(dg, og, hg, ot, hr, rng, 3, 40, 1500, 4000, extra_dels=[(10, 11), (700, 701), (2500, 2501)])


def test_hub_subtree_with_chunked_rows(dg):
    """The invalidated subtree holds a hub whose row is cut into chunks (light rows of 8 edges)."""
    _say("[hub] n=4000")
    rng = np.random.default_rng(6)
    n = 4000
    src = np.concatenate([[0], np.full(3000, 1), rng.integers(0, n, 20000)])
    dst = np.concatenate([[1], np.arange(2, 3002), rng.integers(0, n, 20000)])
    w = np.concatenate([[1], np.ones(3000, np.int64), rng.integers(20, 60, 20000)])
    og = oracle.build_from_edges(n, src, dst, w.reshape(-1, 1))
    hg = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    ot, _ = oracle.compute(og, 0)
    hr = dg.sssp.compute(hg, 0)
    hr.set_tuning(dg.sssp.Tuning(light_degree=8, chunk_edges=64))
    _stream(dg, og, hg, ot, hr, rng, 2, 50, 20, 60, extra_dels=[(0, 1)])


def test_imported_tree_updates_like_the_engine_tree(dg):
    """A validated imported tree is updated by pointer jumping, the next batch by the top-down
    walk: both equal the oracle."""
    _say("[imported] n=3000")
    rng = np.random.default_rng(8)
    n, m = 3000, 30000
    src, dst, w = rng.integers(0, n, m), rng.integers(0, n, m), rng.integers(1, 40, m)
    og = oracle.build_from_edges(n, src, dst, w.reshape(-1, 1))
    hg = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    ot, _ = oracle.compute(og, 7)
    hr = dg.sssp.Result.from_arrays(hg, 7, ot.dist, ot.parent)
    _stream(dg, og, hg, ot, hr, rng, 3, 200, 1, 40)


def test_barrier_timeout_fails_cleanly(dg):
    """A grid barrier whose wait limit is zero fails the run with a device error (every workgroup
    leaves the kernel); the result is poisoned and the device stays usable."""
    _say("[timeout] n=5000")
    rng = np.random.default_rng(9)
    n, m = 5000, 50000
    src, dst, w = rng.integers(0, n, m), rng.integers(0, n, m), rng.integers(1, 40, m)
    og = oracle.build_from_edges(n, src, dst, w.reshape(-1, 1))
    hg = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    ot, _ = oracle.compute(og, 0)
    hr = dg.sssp.compute(hg, 0)
    hr.set_tuning(dg.sssp.Tuning(barrier_timeout_us=-1))
    rows = np.repeat(np.arange(og.n), np.diff(og.row_ptr))
    pick = rng.integers(0, og.m, 300)
    ins = (rng.integers(0, n, 300), rng.integers(0, n, 300), rng.integers(1, 40, 300))
    dels = (rows[pick], og.col[pick].copy())
    og, _, _ = oracle.update(og, ot, ins[0], ins[1], ins[2].reshape(-1, 1), dels[0], dels[1])
    with pytest.raises(dg.DeviceError, match="grid barrier"):
        dg.sssp.update(hg, dg.EdgeBatch(insert=ins, delete=dels), hr)
    with pytest.raises(dg.StaleResultError):
        hr.to_numpy()
    # the next run on the shared workspace starts clean
    fresh = dg.sssp.compute(hg, 0)
    d, p = fresh.to_numpy()
    assert np.array_equal(d, ot.dist) and np.array_equal(p, ot.parent)
    _stream(dg, og, hg, ot, fresh, rng, 2, 200, 1, 40)


def test_merges_planned_on_the_device_and_again_on_the_host(dg):
    """A small batch's merges are planned and run on the device with no read-back of the plan
    (host_plans 0); a batch whose full merge needs staging the graph does not have yet is held
    back on the device (the merges and the engine run queued behind them do nothing) and planned
    again on the host (host_plans >= 1). Trees, counts and rows equal the oracle's either way."""
    _say("[speculative merges] n=3000")
    rng = np.random.default_rng(31)
    n = 3000
    src = np.concatenate([np.zeros(200, np.int64), rng.integers(1, n, 30000)])
    dst = np.concatenate([rng.permutation(np.arange(1, n))[:200], rng.integers(0, n, 30000)])
    w = rng.integers(1, 60, len(src))
    og = oracle.build_from_edges(n, src, dst, w.reshape(-1, 1))
    hg = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    ot, _ = oracle.compute(og, 0)
    hr = dg.sssp.compute(hg, 0)
    plans = []
    for b in range(4):
        rows = np.repeat(np.arange(og.n), np.diff(og.row_ptr))
        pick = rng.choice(og.m, 20, replace=False)
        ins_s, ins_d = rng.integers(1, n, 20), rng.integers(0, n, 20)
        if b % 2 == 1:  # vertex 0's row grows past its tail limit: a full merge through staging
            ins_s = np.concatenate([ins_s, np.zeros(300, np.int64)])
            ins_d = np.concatenate([ins_d, rng.integers(1, n, 300)])
        ins = (ins_s, ins_d, rng.integers(1, 60, len(ins_s)))
        dels = (rows[pick], og.col[pick].copy())
        og, a, c = oracle.update(og, ot, ins[0], ins[1], ins[2].reshape(-1, 1), dels[0], dels[1])
        st = dg.sssp.update(hg, dg.EdgeBatch(insert=ins, delete=dels), hr)
        d, p = hr.to_numpy()
        assert np.array_equal(d, ot.dist) and np.array_equal(p, ot.parent), f"batch {b}"
        assert (st.invalidated, st.affected) == (c.invalidated, c.affected), f"batch {b}"
        _assert_same_rows(hg, og)
        plans.append(st.batch.host_plans)
    _say(f"[speculative merges] host plans per batch {plans}")
    assert plans[0] == 0, "a small batch of tail merges reads no plan back"
    assert plans[1] >= 1, "the first full merge (no staging yet) is planned again on the host"


def test_long_rows_with_many_operations(dg):
    """Rows longer than a wave's share and batches with more operations on one row than the merge
    stages in LDS (both directions: vertex 0's out-row and vertex 1's in-row)."""
    _say("[long rows] n=12000")
    rng = np.random.default_rng(12)
    n = 12000
    hub_out = rng.permutation(np.arange(2, n))[:6000]
    hub_in = rng.permutation(np.arange(2, n))[:6000]
    src = np.concatenate([np.zeros(6000, np.int64), hub_in, rng.integers(0, n, 30000)])
    dst = np.concatenate([hub_out, np.ones(6000, np.int64), rng.integers(0, n, 30000)])
    w = rng.integers(1, 100, len(src))
    og = oracle.build_from_edges(n, src, dst, w.reshape(-1, 1))
    hg = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    _assert_same_rows(hg, og)
    ot, _ = oracle.compute(og, 0)
    hr = dg.sssp.compute(hg, 0)
    for b in range(3):
        dels_out = rng.choice(hub_out, 1500, replace=False)
        dels_in = rng.choice(hub_in, 1500, replace=False)
        ins_out = rng.integers(2, n, 1500)
        ins_in = rng.integers(2, n, 1500)
        ins = (np.concatenate([np.zeros(1500, np.int64), ins_in]),
               np.concatenate([ins_out, np.ones(1500, np.int64)]), rng.integers(1, 100, 3000))
        dels = (np.concatenate([np.zeros(1500, np.int64), dels_in]),
                np.concatenate([dels_out, np.ones(1500, np.int64)]))
        og, a, c = oracle.update(og, ot, ins[0], ins[1], ins[2].reshape(-1, 1), dels[0], dels[1])
        st = dg.sssp.update(hg, dg.EdgeBatch(insert=ins, delete=dels), hr)
        d, p = hr.to_numpy()
        assert np.array_equal(d, ot.dist) and np.array_equal(p, ot.parent), f"batch {b}"
        assert st.invalidated == c.invalidated and st.affected == c.affected
        for f in ("inserted_edges", "updated_edges", "deleted_edges", "ignored_deletions"):
            assert getattr(st.batch, f) == a.summary[f], f
        _assert_same_rows(hg, og)


@pytest.mark.parametrize("hub_is_source", [True, False])
def test_one_store_per_merge_mode(dg, hub_is_source):
    """A batch whose operations all touch one hub: the hub's row (out-row when it is the source of
    every operation, else in-row) takes a workgroup merge while every other touched row takes a
    lane merge in the other store, so each merge launch has rows of one store only."""
    _say(f"[one store per mode] hub_is_source={hub_is_source}")
    rng = np.random.default_rng(21)
    n = 5000
    src = rng.integers(0, n, 40000)
    dst = rng.integers(0, n, 40000)
    w = rng.integers(1, 50, len(src))
    og = oracle.build_from_edges(n, src, dst, w.reshape(-1, 1))
    hg = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    ot, _ = oracle.compute(og, 0)
    hr = dg.sssp.compute(hg, 0)
    hub = 7
    for b, size in enumerate((3000, 40, 2500)):
        others = rng.permutation(np.arange(n))[:size]
        others = others[others != hub]
        hubs = np.full(len(others), hub, np.int64)
        ins_src, ins_dst = (hubs, others) if hub_is_source else (others, hubs)
        ins_w = rng.integers(1, 50, len(others))
        empty = np.zeros(0, np.int64)
        og, a, c = oracle.update(og, ot, ins_src, ins_dst, ins_w.reshape(-1, 1), empty, empty)
        st = dg.sssp.update(hg, dg.EdgeBatch(insert=(ins_src, ins_dst, ins_w)), hr)
        d, p = hr.to_numpy()
        assert np.array_equal(d, ot.dist) and np.array_equal(p, ot.parent), f"batch {b}"
        assert st.batch.inserted_edges == a.summary["inserted_edges"]
        _assert_same_rows(hg, og)


@pytest.mark.parametrize("K,allow_parallel", [(3, True), (5, False), (8, True)])
def test_rows_with_many_weight_columns(dg, K, allow_parallel):
    """Graphs with more weight columns than a tail merge moves per pass (every merge mode: short
    tails by a thread or a wave, long tails by a workgroup, whole rows): after every batch the rows
    and all K columns equal the oracle's, and the tree of the last objective equals Dijkstra's."""
    _say(f"[columns] K={K} parallel={allow_parallel}")
    rng = np.random.default_rng(40 + K)
    n = 6000
    op, gp = _props_pair(dg, allow_parallel=allow_parallel, sorted_rows=not allow_parallel)
    hub = rng.permutation(np.arange(2, n))[:3000]
    src = np.concatenate([np.zeros(3000, np.int64), rng.integers(0, n, 24000)])
    dst = np.concatenate([hub, rng.integers(0, n, 24000)])
    w = rng.integers(1, 60, (len(src), K))
    og = oracle.build_from_edges(n, src, dst, w, op)
    hg = dg.Graph.from_edges(src, dst, w, num_vertices=n, properties=gp)
    _assert_same_rows(hg, og)
    r = dg.sssp.compute(hg, 0, objective=K - 1)
    for b in range(4):
        ni = int(rng.integers(200, 3000))
        nd = int(rng.integers(200, 3000))
        # many operations on the hub's row (a long tail), the rest spread thin (short tails)
        ins_s = np.concatenate([np.zeros(ni // 3, np.int64), rng.integers(0, n, ni - ni // 3)])
        ins_d = rng.integers(1, n, ni)
        ins_w = rng.integers(1, 60, (ni, K))
        rows = np.repeat(np.arange(og.n), np.diff(og.row_ptr))
        pick = rng.integers(0, og.m, nd)
        del_s, del_d = rows[pick], og.col[pick]
        og = oracle.apply_batch(og, ins_s, ins_d, ins_w, del_s, del_d, op).g
        st = dg.sssp.update(hg, dg.EdgeBatch(insert=(ins_s, ins_d, ins_w), delete=(del_s, del_d)), r)
        _assert_same_rows(hg, og)
        d, p = oracle.dijkstra(og, 0, K - 1)
        assert np.array_equal(r.distances, d) and np.array_equal(r.parents, p), f"batch {b}"
        assert st.batch.num_vertices_after == og.n


def test_grid_with_deletions_on_the_tree(dg):
    """The high-diameter configuration at a size the oracle runs in a moment: a 128 x 128 grid
    (4-neighbour arcs both ways, 0.1 % shortcuts, w ~ U[1, 1000], source at the corner) and
    batches whose deletions hit the current shortest-path tree (deep subtrees)."""
    from dyng_amd import generators as G

    side = 128
    _say(f"[grid] side={side}")
    src, dst, w = G.grid(side, 0.001, 1000, G.GRAPH_SEED ^ 2)
    n = side * side
    og = oracle.build_from_edges(n, src, dst, w.reshape(-1, 1))
    hg = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    ot, _ = oracle.compute(og, 0)
    hr = dg.sssp.compute(hg, 0)
    d, p = hr.to_numpy()
    assert np.array_equal(d, ot.dist) and np.array_equal(p, ot.parent)
    for b in range(4):
        ins_s, ins_d, ins_w, del_s, del_d = G.tree_batch(side, ot.parent, 500, 400, 100, 1000,
                                                         G.BATCH_SEED ^ 2, b)
        og, a, c = oracle.update(og, ot, ins_s, ins_d, ins_w.reshape(-1, 1), del_s, del_d)
        st = dg.sssp.update(hg, dg.EdgeBatch(insert=(ins_s, ins_d, ins_w), delete=(del_s, del_d)), hr)
        d, p = hr.to_numpy()
        assert np.array_equal(d, ot.dist), f"distances, batch {b}"
        assert np.array_equal(p, ot.parent), f"parents, batch {b}"
        assert st.invalidated == c.invalidated and st.affected == c.affected
        assert st.invalidated > 0


# ------------------------------------------------------------------------------------------------
# Batched multi-instance repair, repair after apply, the hipGraph launch, the autotuner
# ------------------------------------------------------------------------------------------------

def _rmat_instance(rng, n, wmax):
    """A small RMAT-like instance (a=0.57, b=0.19, c=0.19) with average degree 16."""
    m = 16 * n
    bits = max(1, int(np.ceil(np.log2(n))))
    u = np.zeros(m, np.int64)
    v = np.zeros(m, np.int64)
    for _ in range(bits):
        r = rng.random(m)
        u = 2 * u + (r >= 0.76)
        v = 2 * v + (((r >= 0.57) & (r < 0.76)) | (r >= 0.95))
    u %= n
    v %= n
    keep = u != v
    return u[keep], v[keep], rng.integers(1, wmax + 1, int(keep.sum()))


def _instance_batch(rng, og, n, frac, wmax):
    rows = np.repeat(np.arange(og.n), np.diff(og.row_ptr))
    k = max(1, int(frac * og.m) // 2)
    pick = rng.choice(og.m, min(k, og.m), replace=False)
    ins = (rng.integers(0, n, k), rng.integers(0, n, k), rng.integers(1, wmax + 1, k))
    return ins[0], ins[1], ins[2], rows[pick], og.col[pick].copy()


def test_batched_instances_match_each_instance_alone(dg):
    """Independent RMAT graphs of uneven sizes as one device graph: one compute and one update
    launch for all, each instance's tree equal to the oracle on the instance alone."""
    _say("[batched] 12 instances")
    rng = np.random.default_rng(21)
    sizes = [int(2 ** rng.uniform(6, 12)) for _ in range(12)]
    graphs, oracles, sources = [], [], []
    for n in sizes:
        s, d, w = _rmat_instance(rng, n, 100)
        graphs.append((n, s, d, w))
        og = oracle.build_from_edges(n, s, d, w.reshape(-1, 1))
        src = int(rng.integers(0, n))
        ot, _ = oracle.compute(og, src)
        oracles.append([og, ot])
        sources.append(src)
    inst = dg.batched.Instances(graphs)
    r = inst.compute(sources)
    arrays = r.to_numpy()
    for i, (og, ot) in enumerate(oracles):
        d, p = inst.tree(r, i, arrays)
        assert np.array_equal(d, ot.dist) and np.array_equal(p, ot.parent), f"instance {i} compute"
    for b in range(3):
        batches = []
        for i, n in enumerate(sizes):
            x = _instance_batch(rng, oracles[i][0], n, 0.01, 100)
            batches.append(x)
            oracles[i][0], _, _ = oracle.update(oracles[i][0], oracles[i][1], x[0], x[1], x[2].reshape(-1, 1),
                                                x[3], x[4])
        inst.update(batches, r)
        arrays = r.to_numpy()
        for i, (og, ot) in enumerate(oracles):
            d, p = inst.tree(r, i, arrays)
            assert np.array_equal(d, ot.dist) and np.array_equal(p, ot.parent), f"instance {i} batch {b}"


def test_repair_after_apply_equals_update(dg):
    rng = np.random.default_rng(22)
    n, m = 3000, 30000
    src, dst, w = rng.integers(0, n, m), rng.integers(0, n, m), rng.integers(1, 50, m)
    g1 = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    g2 = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    r1, r2 = dg.sssp.compute(g1, 3), dg.sssp.compute(g2, 3)
    og = oracle.build_from_edges(n, src, dst, w.reshape(-1, 1))
    ot, _ = oracle.compute(og, 3)
    for b in range(3):
        x = _instance_batch(rng, og, n + 2, 0.02, 50)  # a little vertex growth too
        og, _, c = oracle.update(og, ot, x[0], x[1], x[2].reshape(-1, 1), x[3], x[4])
        batch = dg.EdgeBatch(insert=(x[0], x[1], x[2]), delete=(x[3], x[4]))
        st1 = dg.sssp.update(g1, batch, r1)
        g2.apply(batch)
        with pytest.raises(dg.StaleResultError):
            dg.sssp.update(g2, dg.EdgeBatch(), r2)  # one apply behind: update refuses
        st2 = dg.sssp.repair(g2, r2)
        for r, st in ((r1, st1), (r2, st2)):
            d, p = r.to_numpy()
            assert np.array_equal(d, ot.dist) and np.array_equal(p, ot.parent), f"batch {b}"
            assert st.invalidated == c.invalidated and st.affected == c.affected
    with pytest.raises(dg.StaleResultError):
        dg.sssp.repair(g2, r2)  # already current


def test_results_sharing_the_workspace_interleave(dg):
    """Two results of one graph repaired in turn, and a clone: each run finds the workspace holding
    the other tree's words (or its own), so the engine must tell whether it may skip packing; every
    tree stays equal to the oracle's canonical tree after every batch."""
    rng = np.random.default_rng(23)
    n, m = 4000, 40000
    src, dst, w = rng.integers(0, n, m), rng.integers(0, n, m), rng.integers(1, 60, m)
    g = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    og = oracle.build_from_edges(n, src, dst, w.reshape(-1, 1))
    r1, r2 = dg.sssp.compute(g, 5), dg.sssp.compute(g, 77)
    for b in range(5):
        x = _instance_batch(rng, og, n, 0.02, 60)
        og, _, _ = oracle.update(og, oracle.compute(og, 5)[0], x[0], x[1], x[2].reshape(-1, 1), x[3], x[4])
        g.apply(dg.EdgeBatch(insert=(x[0], x[1], x[2]), delete=(x[3], x[4])))
        order = (r1, r2) if b % 2 == 0 else (r2, r1)
        for r in order:
            dg.sssp.repair(g, r)
        if b == 2:
            r1 = r1.clone()  # a copy of the tree whose words the workspace may hold
        for r, s0 in ((r1, 5), (r2, 77)):
            ot, _ = oracle.compute(og, s0)
            d, p = r.to_numpy()
            assert np.array_equal(d, ot.dist) and np.array_equal(p, ot.parent), f"batch {b}, source {s0}"


def test_graph_launch_replays_match_plain_launches(dg):
    """The engine's launch as a hipGraph, replayed for every update (with other eager work and
    synchronizations in between), gives the oracle's tree every time."""
    rng = np.random.default_rng(23)
    n, m = 4000, 48000
    src, dst, w = rng.integers(0, n, m), rng.integers(0, n, m), rng.integers(1, 60, m)
    og = oracle.build_from_edges(n, src, dst, w.reshape(-1, 1))
    ot, _ = oracle.compute(og, 0)
    hg = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    hr = dg.sssp.compute(hg, 0)
    hr.set_tuning(dg.sssp.Tuning(launch="graph"))
    other = dg.Graph.from_edges(src, dst, w, num_vertices=n)  # eager work between the replays
    ro = dg.sssp.compute(other, 1)
    for b in range(5):
        x = _instance_batch(rng, og, n, 0.02, 60)
        og, _, c = oracle.update(og, ot, x[0], x[1], x[2].reshape(-1, 1), x[3], x[4])
        st = dg.sssp.update(hg, dg.EdgeBatch(insert=(x[0], x[1], x[2]), delete=(x[3], x[4])), hr)
        d, p = hr.to_numpy()
        assert np.array_equal(d, ot.dist) and np.array_equal(p, ot.parent), f"replay {b}"
        assert st.invalidated == c.invalidated and st.affected == c.affected
        dg.sssp.update(other, dg.EdgeBatch(insert=(x[1], x[0], x[2])), ro)
        dg.synchronize()


def test_autotune_keeps_the_result(dg):
    rng = np.random.default_rng(24)
    n, m = 5000, 80000
    src, dst, w = rng.integers(0, n, m), rng.integers(0, n, m), rng.integers(1, 100, m)
    og = oracle.build_from_edges(n, src, dst, w.reshape(-1, 1))
    ot, _ = oracle.compute(og, 0)
    hg = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    hr = dg.sssp.compute(hg, 0)
    x = _instance_batch(rng, og, n, 0.02, 100)
    og, _, _ = oracle.update(og, ot, x[0], x[1], x[2].reshape(-1, 1), x[3], x[4])
    rep = dg.batched.autotune(hg, dg.EdgeBatch(insert=(x[0], x[1], x[2]), delete=(x[3], x[4])), hr)
    assert len(rep.table) == len(dg.batched.candidate_tunings(1)) and rep.engine_ms > 0
    d, p = hr.to_numpy()
    assert np.array_equal(d, ot.dist) and np.array_equal(p, ot.parent)
    # the tuned schedule keeps giving the oracle's tree
    x = _instance_batch(rng, og, n, 0.02, 100)
    og, _, _ = oracle.update(og, ot, x[0], x[1], x[2].reshape(-1, 1), x[3], x[4])
    dg.sssp.update(hg, dg.EdgeBatch(insert=(x[0], x[1], x[2]), delete=(x[3], x[4])), hr)
    d, p = hr.to_numpy()
    assert np.array_equal(d, ot.dist) and np.array_equal(p, ot.parent)
