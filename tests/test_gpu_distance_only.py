"""The HIP engine's distance-only mode, the 47-bit packing boundary, canonicalization on import and
the non-canonical tie case, against the oracle and the brute-force tree (tests/cases.py).

The reference runs these shapes in cpp/tests/algorithms/sssp/sssp_test.cpp:239-349 and
sssp_cuda_test.cpp:501-617: weights of 2 * 10^9 (or random up to 2^31 - 1) on grids of more than
2^16 vertices, where (n - 1) * max weight does not fit beside the parent ids, so the words hold
distances only and the parents are recovered after the search (lowest-id tight in-neighbour).
Needs an MI355X: every test here is marked gpu.
"""

from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
from tests import cases

pytestmark = pytest.mark.gpu

W_TIE = 2_000_000_000


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dg():
    import dyng_amd

    dyng_amd._lib.lib()  # loud failure if the HIP library is missing
    assert dyng_amd.device_count() > 0, "no HIP device visible"
    return dyng_amd


def _pair(dg, n, src, dst, w):
    w = np.asarray(w, np.int64).reshape(len(src), -1)
    props = dg.GraphProperties.mosp_compatible()
    og = oracle.build_from_edges(n, src, dst, w, oracle.Props.mosp_compatible())
    hg = dg.Graph.from_edges(np.asarray(src), np.asarray(dst), w, num_vertices=n, properties=props)
    return og, hg


def _same(r, d, p, what):
    hd, hp = r.to_numpy()
    assert np.array_equal(hd, d), f"{what}: distances"
    assert np.array_equal(hp, p), f"{what}: parents"


def test_distance_only_canonicalizes_non_canonical_input(dg):
    # sssp_test.cpp:239-273
    side = 272
    _say(f"[distance-only] one-way grid {side}x{side}, w = 2e9, non-canonical import")
    n, s, d, w = cases.grid_one_way(side, W_TIE)
    og, hg = _pair(dg, n, s, d, w)
    canon = dg.sssp.compute(hg, 0)
    bd, bp = cases.brute_tree(n, s, d, w, 0)
    _same(canon, bd, bp, "compute")
    p = bp.copy().reshape(side, side)
    p[1:, 1:] = np.arange(n, dtype=np.int32).reshape(side, side)[1:, 1:] - 1
    p = p.ravel()
    r = dg.sssp.Result.from_arrays(hg, 0, bd, p, canonicalize=False)
    st = dg.sssp.update(hg, dg.EdgeBatch(delete=([0], [1])), r)
    assert not st.packed_parents
    ot = oracle.Tree(0, bd.copy(), p.copy())
    og2, _, c = oracle.update(og, ot, [], [], [], [0], [1], props=oracle.Props.mosp_compatible())
    assert not c.packed_parents
    _same(r, ot.dist, ot.parent, "update")
    assert st.invalidated == c.invalidated and st.affected == c.affected
    fresh = dg.sssp.compute(hg, 0)
    _same(r, *fresh.to_numpy(), "update vs compute")


def test_distance_only_keeps_canonical_parents(dg):
    # sssp_test.cpp:275-315
    side = 320
    _say(f"[distance-only] two-way grid {side}x{side}, w = 2e9")
    n, s, d = cases.grid_two_way(side)
    w = np.full(len(s), W_TIE, np.int64)
    og, hg = _pair(dg, n, s, d, w)
    r = dg.sssp.compute(hg, 0)
    ot, _ = oracle.compute(og, 0)
    _same(r, ot.dist, ot.parent, "compute")
    v = np.arange(1, n, 97)
    og2, _, c = oracle.update(og, ot, [0], [n - 1], [[W_TIE]], v - 1, v, props=oracle.Props.mosp_compatible())
    st = dg.sssp.update(hg, dg.EdgeBatch(insert=([0], [n - 1], [W_TIE]), delete=(v - 1, v)), r)
    assert not st.packed_parents
    _same(r, ot.dist, ot.parent, "update")
    assert st.invalidated == c.invalidated and st.affected == c.affected
    hd, hp = r.to_numpy()
    assert oracle.check_tree(og2, 0, hd, hp, require_canonical=True)["ok"]
    _same(dg.sssp.compute(hg, 0), hd, hp, "update vs compute")


def test_large_weights_use_the_distance_only_mode(dg):
    # sssp_cuda_test.cpp:559-617: two objectives (2 * 10^9 everywhere; random in [1, 2^31 - 1]),
    # three rounds each of ~1000 tree-edge deletions and ~1000 insertions, against the OpenMP
    # engine of the oracle
    side = 320
    n, s, d = cases.grid_two_way(side)
    rng = np.random.default_rng(20260927)
    w = np.stack([np.full(len(s), W_TIE, np.int64), rng.integers(1, 2**31, len(s))], 1)
    for k in (0, 1):
        _say(f"[distance-only] large weights, objective {k}")
        og, hg = _pair(dg, n, s, d, w)
        r = dg.sssp.compute(hg, 0, objective=k)
        ot, _ = oracle.compute(og, 0, k, engine="omp", threads=4)
        _same(r, ot.dist, ot.parent, "compute")
        for rnd in range(3):
            v = rng.integers(1, n, 1000)
            v = v[ot.parent[v] >= 0]
            a, b = rng.integers(0, n, 1000), rng.integers(0, n, 1000)
            keep = a != b
            iw = np.stack([np.full(keep.sum(), W_TIE, np.int64), rng.integers(1, 2**31, keep.sum())], 1)
            dels = (ot.parent[v].copy(), v)
            og, _, c = oracle.update(og, ot, a[keep], b[keep], iw, *dels, props=oracle.Props.mosp_compatible(),
                                     engine="omp", threads=4)
            st = dg.sssp.update(hg, dg.EdgeBatch(insert=(a[keep], b[keep], iw), delete=dels), r)
            assert not st.packed_parents and not c.packed_parents
            _same(r, ot.dist, ot.parent, f"objective {k} round {rnd}")
            assert (st.invalidated, st.affected) == (c.invalidated, c.affected), f"objective {k} round {rnd}"
            hd, hp = r.to_numpy()
            assert oracle.check_tree(og, 0, hd, hp, k=k, require_canonical=True)["ok"]


@pytest.mark.parametrize("scenario", ["pull", "push", "from-scratch"])
def test_packing_boundary_of_47_distance_bits(dg, scenario):
    # sssp_cuda_test.cpp:501-553: the device packs ((n - 1) W fits in 47 bits); the back edge
    # n-1 -> 1 offers n W, which must never be packed
    n, W = (1 << 17) - 1, (1 << 30) + (1 << 14)
    _say(f"[packing boundary] path n={n}, W={W}, {scenario}")
    back = scenario != "pull"
    og, hg = _pair(dg, n, *cases.path(n, W, back))
    r = dg.sssp.compute(hg, 0)
    vid = np.arange(n, dtype=np.int64)
    _same(r, vid * W, (vid - 1).astype(np.int32), "compute")
    if scenario == "from-scratch":
        return
    ins = (n - 1, 1, W) if scenario == "pull" else (n - 2, n - 1, W - 1)
    ot = oracle.Tree(0, vid * W, (vid - 1).astype(np.int32))
    _, _, c = oracle.update(og, ot, [ins[0]], [ins[1]], [[ins[2]]], [], [], props=oracle.Props.mosp_compatible())
    st = dg.sssp.update(hg, dg.EdgeBatch(insert=([ins[0]], [ins[1]], [ins[2]])), r)
    assert st.packed_parents
    _same(r, ot.dist, ot.parent, scenario)
    assert st.affected == c.affected == (1 if scenario == "push" else 0)


def test_canonicalize_on_import(dg):
    # sssp.cpp:1153-1156: from_arrays(canonicalize=True) gives every vertex its lowest-id tight
    # parent; the planted trees are valid but not canonical
    rng = np.random.default_rng(31)
    planted = 0
    for it in range(6):
        n = int(rng.integers(500, 3000))
        m = int(rng.integers(3 * n, 8 * n))
        s, d, w = rng.integers(0, n, m), rng.integers(0, n, m), rng.integers(1, 4, m)
        _say(f"[canonicalize] n={n} m={m}")
        og = oracle.build_from_edges(n, s, d, w.reshape(-1, 1), oracle.Props(allow_parallel=True))
        props = dg.GraphProperties(parallel_edges="allow")
        hg = dg.Graph.from_edges(s, d, w, num_vertices=n, properties=props)
        dist, par = oracle.dijkstra(og, 0)
        rows = np.repeat(np.arange(og.n), np.diff(og.row_ptr))
        p = cases.random_tight_parents(rng, n, rows, og.col, og.w[0], dist, par, 0)
        planted += int((p != par).sum())
        r = dg.sssp.Result.from_arrays(hg, 0, dist, p, canonicalize=True)
        _same(r, dist, oracle.canonicalize(og, 0, dist, p), "canonicalized")
        _same(r, dist, par, "canonicalized vs Dijkstra")
        # and imported as is, the parents stay as planted
        r2 = dg.sssp.Result.from_arrays(hg, 0, dist, p, canonicalize=False)
        _same(r2, dist, p, "imported as is")
    assert planted > 100


def test_non_canonical_input_keeps_tie_parents_like_mosp(dg):
    # sssp_test.cpp:317-349
    _say("[ties] non-canonical input tree")
    edges = np.array([(0, 1, 1), (0, 2, 1), (1, 3, 1), (2, 3, 1), (0, 4, 5), (4, 3, 1)], np.int64)

    def graph():
        return dg.Graph.from_edges(edges[:, 0], edges[:, 1], edges[:, 2], num_vertices=5,
                                   properties=dg.GraphProperties.mosp_compatible())

    d0, p0 = np.array([0, 1, 1, 2, 5]), np.array([-1, 0, 0, 2, 0])
    g = graph()
    r = dg.sssp.Result.from_arrays(g, 0, d0, p0, canonicalize=False)
    st = dg.sssp.update(g, dg.EdgeBatch(insert=([0], [4], [1])), r)
    d, p = r.to_numpy()
    assert d.tolist() == [0, 1, 1, 2, 1] and p.tolist() == [-1, 0, 0, 2, 0]
    assert st.affected == 1
    fresh = dg.sssp.compute(g, 0)
    fd, fp = fresh.to_numpy()
    assert fd.tolist() == d.tolist() and fp.tolist() == [-1, 0, 0, 1, 0]
    g2 = graph()
    c = dg.sssp.Result.from_arrays(g2, 0, d0, p0, canonicalize=True)
    dg.sssp.update(g2, dg.EdgeBatch(insert=([0], [4], [1])), c)
    assert c.to_numpy()[1].tolist() == fp.tolist()
