"""CPU tests of the hyper_sssp oracles (oracle/hyper.py, oracle/hypergraph_model.py) and of the
host side of dyng_amd.hyper_sssp in libdyng_hip.so: the hypergraph container (ids, skip rules,
counts) and the edge batch that carries a result's device line graph from one hypergraph version
to the next. The library's host calls run without a GPU.

Pins of the oracle: the reference's worked example and test expectations
(python/dyng/hyper_sssp.py docstring; python/tests/test_hyper_sssp.py:63-82,104-112),
Bellman-Ford over the line graph on random hypergraphs, and planted mistakes that these checks
catch. Pins of the container: the reference's own expectations (python/tests/test_hypergraph.py:
50-58,60-70,195-241,244-258), and the reference test model's validity rules for every id it gives
(python/dyng/testing/oracles.py:129-235, restated in `_model_apply`).
"""

from __future__ import annotations

import numpy as np
import pytest

import dyng_amd as dg
from dyng_amd import hyper_sssp as hs
from oracle import hyper as oh
from oracle import hypergraph_model as hm

INF = oh.INF
LISTS = [[0, 1], [1, 2], [2, 3], [0, 3], [4, 5]]
WEIGHTS = [2, 3, 1, 9, 4]


# ---- the oracle ----

def test_oracle_worked_example():
    """python/dyng/hyper_sssp.py docstring and test_compute_example (test_hyper_sssp.py:63-82)."""
    d, p = oh.reference(LISTS, WEIGHTS, 0)
    assert d == [2, 5, 6, 9, INF] and p == [-1, 0, 1, -1, -1]
    assert oh.vertex_distances(LISTS, d, 6, 0) == [0, 2, 5, 6, INF, INF]
    d4, p4 = oh.reference(LISTS[:4], WEIGHTS[:4], 0)
    assert d4 == [2, 5, 6, 9] and p4 == [-1, 0, 1, -1]
    # the docstring's update: vertex 3 joins hyperedge 0, so hyperedge 2 is entered from 0
    rows = [[0, 1, 3], [1, 2], [2, 3], [0, 3]]
    d, p = oh.reference(rows, WEIGHTS[:4], 0)
    assert d == [2, 5, 3, 9]
    deleted, inserted = oh.line_graph_delta(LISTS[:4], rows)
    assert (len(inserted), len(deleted)) == (1, 0) and inserted == [(0, 2)]


def _bellman_ford(rows, weights, source):
    m = len(rows)
    adj = oh.line_graph(rows)
    d = [INF] * m
    for h, r in enumerate(rows):
        if source in r:
            d[h] = weights[h]
    for _ in range(m + 1):
        for a in range(m):
            if d[a] >= INF:
                continue
            for b in adj[a]:
                if d[a] + weights[b] < d[b]:
                    d[b] = d[a] + weights[b]
    p = [-1] * m
    for h in range(m):
        if d[h] >= INF or (source in rows[h] and d[h] == weights[h]):
            continue
        tight = [g for g in adj[h] if d[g] < INF and d[g] + weights[h] == d[h]]
        p[h] = min(tight)
    return d, p


def _random_rows(rng, n, m, kmax, dead=0.0):
    rows = []
    for _ in range(m):
        if rng.random() < dead:
            rows.append([])
        else:
            rows.append(sorted(set(rng.integers(0, n, int(rng.integers(1, kmax + 1))).tolist())))
    return rows


@pytest.mark.parametrize("seed", range(12))
def test_oracle_against_bellman_ford(seed):
    rng = np.random.default_rng(seed)
    n, m = int(rng.integers(3, 25)), int(rng.integers(1, 30))
    rows = _random_rows(rng, n, m, 4, dead=0.15)
    w = rng.integers(1, 6, m).tolist()  # small weights: many equal-distance ties
    for s in range(min(n, 5)):
        assert oh.reference(rows, w, s) == _bellman_ford(rows, w, s)


def test_planted_mistakes_are_caught():
    # seeding the source's hyperedges at 0 instead of w(h)
    def seeded_at_zero(rows, weights, source):
        d, p = oh.reference(rows, [0 if source in r else x for r, x in zip(rows, weights)], source)
        return d, p

    assert seeded_at_zero(LISTS, WEIGHTS, 0)[0] != [2, 5, 6, 9, INF]
    # the highest-id tight parent instead of the lowest: a tie case
    rows, w = [[0, 1], [0, 2], [1, 2, 3]], [1, 1, 5]
    d, p = oh.reference(rows, w, 0)
    assert d[2] == 6 and p[2] == 0  # both 0 and 1 are tight; the lowest id wins
    lg = oh.line_graph(rows)
    assert max(g for g in lg[2] if d[g] + w[2] == d[2]) != p[2]


# ---- the container ----

def test_ids_of_from_lists_and_an_empty_hypergraph():
    """test_hypergraph.py:50-58 and 60-70."""
    h = dg.Hypergraph.from_lists([[0, 1, 2], [1, 3], [4]])
    assert (h.num_vertices, h.num_hyperedges, h.hyperedge_id_bound, h.version) == (5, 3, 3, 0)
    assert h.rows() == [[0, 1, 2], [1, 3], [4]] and not h.weighted
    h = dg.Hypergraph.from_lists([], num_vertices=4)
    s = h.apply(dg.HyperedgeBatch(insert_hyperedges=[[0, 3], [1]]))
    assert s.inserted_ids.tolist() == [0, 1] and h.version == 1
    h = dg.Hypergraph.from_lists([[0, 1], [1, 2]], [5, 7])
    s = h.apply(dg.HyperedgeBatch([0], [[0, 2]], [9]))
    assert h.weights[s.inserted_ids[0]] == 9
    with pytest.raises(dg.InvalidArgumentError, match="weight"):
        h.apply(dg.HyperedgeBatch(insert_hyperedges=[[0]]))
    with pytest.raises(dg.InvalidArgumentError):
        dg.Hypergraph.from_lists([[0]], [0])


@pytest.mark.parametrize(("reuse", "want"), [("erase_first", 1), ("insert_first", 3)])
def test_the_id_reuse_example(reuse, want):
    """test_hypergraph.py:229-241 (ADR 0041 section 9)."""
    h = dg.Hypergraph.from_lists([[0, 1, 2, 3], [4], [5, 6]], num_vertices=9, id_reuse=reuse, capacity="escher_gpu")
    s = h.apply(dg.HyperedgeBatch(delete_hyperedges=[1], insert_hyperedges=[[7, 8]]))
    assert s.inserted_ids.tolist() == [want] and s.deleted_hyperedges == 1


def test_skipped_ops_are_counted():
    """test_hypergraph.py:244-258."""
    h = dg.Hypergraph.from_lists([[0, 1], [1, 2], [2, 3]])
    h.apply(dg.HyperedgeBatch(delete_hyperedges=[2]))
    s = h.apply(dg.HyperedgeBatch(delete_hyperedges=[0, 0, 2, 7, -3], insert_hyperedges=[[3, 3, 1, 1, 0]]))
    assert (s.deleted_hyperedges, s.ignored_deletions, s.dropped_members) == (1, 4, 2)
    assert s.skipped_ops == 6 and s.num_hyperedges_after == h.num_hyperedges == 2
    assert s.id_bound_after == h.hyperedge_id_bound
    assert h.rows()[s.inserted_ids[0]] == [0, 1, 3]


@pytest.mark.parametrize("reuse", ["erase_first", "insert_first"])
def test_incidence_edits(reuse):
    """test_hypergraph.py:195-226: removals (applied, a repeat, applied, the last member, a vertex
    out of range, a dead id, a hyperedge the batch deletes) and additions (applied, a member, a
    hyperedge the batch deletes, an id out of range, a hyperedge the batch deletes)."""
    lists = [[0, 1, 2], [2, 3], [4, 5], [1, 5]]
    h = dg.Hypergraph.from_lists(lists, id_reuse=reuse)
    removals = [(0, 1), (0, 1), (1, 3), (1, 2), (2, 9), (7, 0), (3, 1)]
    additions = [(2, 0), (2, 4), (3, 3), (-1, 0), (3, 1)]
    b = dg.HyperedgeBatch(delete_hyperedges=[3], insert_hyperedges=[[0, 3]],
                          delete_incidences=([e[0] for e in removals], [e[1] for e in removals]),
                          insert_incidences=([e[0] for e in additions], [e[1] for e in additions]))
    assert b.num_incidence_edits == 12 and "incidence_edits=12" in repr(b)
    model = {i: list(r) for i, r in enumerate(lists)}
    s = h.apply(b)
    want = _model_apply(model, 6, reuse, [3], [[0, 3]], s.inserted_ids.tolist(), removals, additions)
    for name, value in want.items():
        assert getattr(s, name) == value, name
    assert (s.deleted_incidences, s.inserted_incidences, s.skipped_ops) == (2, 1, 9)
    assert h.rows() == [model.get(i, []) for i in range(h.hyperedge_id_bound)]
    lg = h.h2h()
    assert [lg.row(i).tolist() for i in range(lg.num_rows)] == oh.line_graph(h.rows())


def _model_apply(rows, n, reuse, deletions, insertions, ids, removals=(), additions=()):
    """The reference test model's apply (python/dyng/testing/oracles.py:166-235): checks that the
    ids are a valid choice and returns the counts a summary must report."""
    before = set(rows)
    deleted, ignored = set(), 0
    for h in deletions:
        if h in before and h not in deleted:
            deleted.add(h)
        else:
            ignored += 1
    new = [sorted(set(h)) for h in insertions]
    dropped = sum(len(h) for h in insertions) - sum(len(r) for r in new)
    assert len(ids) == len(new) and len(set(ids)) == len(ids)
    live_then = before if reuse == "insert_first" else before - deleted
    assert all(i >= 0 and i not in live_then for i in ids)
    for h in deleted:
        del rows[h]
    removed = added = skipped = 0
    for h, v in removals:
        row = rows.get(h) if h not in deleted else None
        if row is None or not 0 <= v < n or v not in row or len(row) == 1:
            skipped += 1
        else:
            row.remove(v)
            removed += 1
    for h, v in additions:
        row = rows.get(h) if h not in deleted else None
        if row is None or not 0 <= v < n or v in row:
            skipped += 1
        else:
            row.append(v)
            row.sort()
            added += 1
    for k, i in enumerate(ids):
        rows[i] = new[k]
    return {"inserted_hyperedges": len(new), "deleted_hyperedges": len(deleted), "ignored_deletions": ignored,
            "dropped_members": dropped, "inserted_incidences": added, "deleted_incidences": removed,
            "skipped_ops": ignored + dropped + skipped, "num_hyperedges_after": len(rows)}


def _random_batch(rng, h, n, weighted=True, size=6):
    bound = max(h.hyperedge_id_bound, 1)
    dels = rng.integers(-1, bound + 1, int(rng.integers(0, size))).tolist()
    ins = [rng.integers(0, n, int(rng.integers(1, 5))).tolist() for _ in range(int(rng.integers(0, size)))]
    w = rng.integers(1, 9, len(ins)).tolist() if weighted else None
    k1, k2 = int(rng.integers(0, size)), int(rng.integers(0, size))
    rem = (rng.integers(0, bound + 1, k1).tolist(), rng.integers(0, n, k1).tolist())
    add = (rng.integers(0, bound + 1, k2).tolist(), rng.integers(0, n, k2).tolist())
    return dels, ins, w, rem, add


@pytest.mark.parametrize("reuse", ["erase_first", "insert_first"])
@pytest.mark.parametrize("seed", range(6))
def test_random_batches_against_the_model(reuse, seed):
    rng = np.random.default_rng(100 + seed)
    n = 14
    lists = [rng.integers(0, n, int(rng.integers(1, 5))).tolist() for _ in range(10)]
    h = dg.Hypergraph.from_lists(lists, rng.integers(1, 9, 10).tolist(), num_vertices=n, id_reuse=reuse)
    model = {i: sorted(set(r)) for i, r in enumerate(lists)}
    for _ in range(12):
        dels, ins, w, rem, add = _random_batch(rng, h, n)
        s = h.apply(dg.HyperedgeBatch(dels, ins, w, delete_incidences=rem, insert_incidences=add))
        want = _model_apply(model, n, reuse, dels, ins, s.inserted_ids.tolist(), list(zip(*rem)), list(zip(*add)))
        for name, value in want.items():
            assert getattr(s, name) == value, name
        assert h.rows() == [model.get(i, []) for i in range(h.hyperedge_id_bound)]
        assert all(not r for r in h.rows()[len(model) and max(model) + 1:])


# ---- the device line graph's edge batch (host side of hyper_sssp.update) ----

def _edges(h, source):
    s, d, w = hs._line_graph_edges(h, source)
    return {(int(a), int(b)): int(x) for a, b, x in zip(s, d, w)}


@pytest.mark.parametrize("reuse", ["erase_first", "insert_first"])
@pytest.mark.parametrize("seed", range(8))
def test_edge_batch_carries_the_line_graph(reuse, seed):
    """Applying the edge batch of an update to the line graph of the old hypergraph (deletions
    first, then upserts) gives the line graph of the new one, weights included, and the pair
    counts equal the oracle's delta."""
    _check_edge_batches(np.random.default_rng(7 + seed), reuse, n=12, m=9, size=6, rounds=10)


@pytest.mark.parametrize("reuse", ["erase_first", "insert_first"])
@pytest.mark.parametrize("seed", range(3))
def test_edge_batch_carries_the_line_graph_mid_size(reuse, seed):
    """The same on 200 hyperedges over 60 vertices (about eight holders per vertex) with batches of
    up to 60 operations of each kind: many touched ids share vertices with each other, so pairs of
    touched ids and ids that are both deleted from and added to a row's neighbours are common."""
    _check_edge_batches(np.random.default_rng(70 + seed), reuse, n=60, m=200, size=60, rounds=8)


def _check_edge_batches(rng, reuse, n, m, size, rounds):
    lists = [rng.integers(0, n, int(rng.integers(1, 5))).tolist() for _ in range(m)]
    h = dg.Hypergraph.from_lists(lists, rng.integers(1, 9, m).tolist(), num_vertices=n, id_reuse=reuse)
    source = int(rng.integers(0, n))
    edges = _edges(h, source)
    assert edges == _brute_edges(h.rows(), h.weights.tolist(), source)
    for _ in range(rounds):
        pre_rows = h.rows()
        dels, ins, w, rem, add = _random_batch(rng, h, n, size=size)
        h.apply(dg.HyperedgeBatch(dels, ins, w, delete_incidences=rem, insert_incidences=add))
        eb, gained, lost = hs._edge_batch(h, source)
        # the library's arrays skip the constructor's checks: its fields are what the constructor builds
        built = dg.EdgeBatch(insert=tuple(a.reshape(-1) for a in eb._ins), delete=eb._del)
        for got, want in zip(eb._ins + eb._del, built._ins + built._del):
            assert got.dtype == want.dtype == np.int32 and got.shape == want.shape and got.flags.c_contiguous
            assert np.array_equal(got, want)
        ins_s, ins_d, ins_w = (np.asarray(a) for a in eb._ins) if eb._ins else ([], [], [])
        del_s, del_d = (np.asarray(a) for a in eb._del) if eb._del else ([], [])
        for a, b in zip(del_s, del_d):
            assert (int(a), int(b)) in edges
            del edges[(int(a), int(b))]
        for a, b, x in zip(ins_s, ins_d, np.asarray(ins_w).reshape(-1)):
            edges[(int(a), int(b))] = int(x)
        assert edges == _edges(h, source) == _brute_edges(h.rows(), h.weights.tolist(), source)
        deleted, inserted = oh.line_graph_delta(pre_rows, h.rows())
        assert (gained, lost) == (len(inserted), len(deleted))


def _brute_edges(rows, weights, source):
    """The device line graph by brute force: (0, h + 1, w(h)) for the hyperedges that hold the
    source, (g + 1, h + 1, w(h)) for the ordered pairs of live hyperedges sharing a vertex."""
    out = {}
    for h, r in enumerate(rows):
        if source in r:
            out[(0, h + 1)] = weights[h]
        for g, q in enumerate(rows):
            if g != h and r and q and set(r) & set(q):
                out[(g + 1, h + 1)] = weights[h]
    return out


# ---- the library's container against the restated rules ----

POLICIES = {
    "dyng_default": hm.CapacityPolicy(),
    "escher_gpu": hm.CapacityPolicy.escher_gpu(),
    # a small payload: compactions, growth and the headroom after a compaction on every batch
    "tight": hm.CapacityPolicy(initial_headroom=0.0, initial_slack=8, compaction_headroom=8),
    "tight_no_compaction": hm.CapacityPolicy(initial_headroom=0.0, initial_slack=8, compact_before_grow=False),
}


def _props_of(p):
    return dg.CapacityPolicy(p.initial_headroom, p.initial_slack, p.min_initial_payload, p.grow_factor,
                             p.compact_before_grow, p.max_payload, p.row_alignment, p.compaction_headroom)


@pytest.mark.parametrize("policy", sorted(POLICIES))
@pytest.mark.parametrize("reuse", ["erase_first", "insert_first"])
@pytest.mark.parametrize("seed", range(4))
def test_container_against_the_restated_rules(policy, reuse, seed):
    """The library's container gives the ids, rows and counts of the restated allocator and apply
    rules, batch after batch, under payload policies that compact and grow often."""
    rng = np.random.default_rng(300 + seed)
    n = 16
    lists = [rng.integers(0, n, int(rng.integers(1, 6))).tolist() for _ in range(12)]
    w = rng.integers(1, 9, 12).tolist()
    pol = POLICIES[policy]
    h = dg.Hypergraph.from_lists(lists, w, num_vertices=n, id_reuse=reuse, capacity=_props_of(pol))
    model = hm.Model(lists, w, n, reuse, pol)
    for _ in range(15):
        dels, ins, iw, rem, add = _random_batch(rng, h, n, size=8)
        s = h.apply(dg.HyperedgeBatch(dels, ins, iw, delete_incidences=rem, insert_incidences=add))
        counts, ids = model.apply(dels, ins, iw, rem, add)
        assert s.inserted_ids.tolist() == ids
        for name, value in counts.items():
            assert getattr(s, name) == value, name
        assert h.rows() == model.rows_by_id()
        assert h.weights.tolist() == [model.w[i] if i < len(model.w) and model.rows_by_id()[i] else 0
                                      for i in range(h.hyperedge_id_bound)]


def test_capacity_error_changes_nothing():
    """A policy that never grows: a batch whose surplus rows do not fit raises CapacityError and
    leaves the hypergraph as it was (the restated allocator raises too)."""
    pol = hm.CapacityPolicy(initial_headroom=0.0, initial_slack=5, grow_factor=0.0)  # 8 + 5 payload slots
    lists = [[0, 1], [1, 2]]
    h = dg.Hypergraph.from_lists(lists, [1, 1], num_vertices=4, capacity=_props_of(pol))
    model = hm.Model(lists, [1, 1], 4, "erase_first", pol)
    big = [[0, 1, 2, 3]] * 3
    with pytest.raises(hm.CapacityError):
        model.apply([], big, [1, 1, 1])
    before = (h.version, h.rows(), h.hyperedge_id_bound)
    with pytest.raises(dg.CapacityError):
        h.apply(dg.HyperedgeBatch(insert_hyperedges=big, insert_weights=[1, 1, 1]))
    assert (h.version, h.rows(), h.hyperedge_id_bound) == before
    s = h.apply(dg.HyperedgeBatch(insert_hyperedges=[[3]], insert_weights=[2]))
    assert s.inserted_ids.tolist() == model.apply([], [[3]], [2])[1]
