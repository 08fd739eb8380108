"""Parity of dyng_amd.hyper_sssp (the line graph on the HIP engine) with the oracle
(oracle/hyper.py) and with the reference's test expectations (python/tests/test_hyper_sssp.py).
Needs an MI355X: every test here is marked gpu.

The bar is bit-exactness: hyperedge distances and canonical parents after compute() and after
every update equal the oracle's on the same hypergraph, an updated result equals compute() on
the new hypergraph, and the line graph's pair counts equal the oracle's delta.
"""

from __future__ import annotations

import numpy as np
import pytest

import dyng_amd as dg
from dyng_amd import hyper_sssp as hs
from oracle import hyper as oh

pytestmark = pytest.mark.gpu

INF = hs.INFINITE_DISTANCE
LISTS = [[0, 1], [1, 2], [2, 3], [0, 3], [4, 5]]
WEIGHTS = [2, 3, 1, 9, 4]


def _say(*a):
    print(*a, flush=True)


def check(h, sp):
    """The result equals the oracle and compute() on the hypergraph."""
    d, p = oh.reference(h.rows(), h.weights.tolist(), sp.source)
    assert sp.hyperedge_distances.tolist() == d
    assert sp.hyperedge_parents.tolist() == p
    again = hs.compute(h, sp.source, options=sp.options)
    assert again.hyperedge_distances.tolist() == d and again.hyperedge_parents.tolist() == p
    assert sp.graph_version == h.version
    assert sp.vertex_distances().tolist() == oh.vertex_distances(h.rows(), d, h.num_vertices, sp.source)


def test_compute_example():
    """test_hyper_sssp.py:63-75 and the module docstring."""
    _say("[hyper_sssp] compute example")
    h = dg.Hypergraph.from_lists(LISTS, WEIGHTS)
    sp = hs.compute(h, source=0)
    assert sp.hyperedge_distances.tolist() == [2, 5, 6, 9, INF]
    assert sp.hyperedge_parents.tolist() == [-1, 0, 1, -1, -1]
    assert sp.source == 0 and sp.graph_version == 0 and sp.space == "device"
    assert sp.hyperedge_distances.dtype == np.int64 and sp.hyperedge_parents.dtype == np.int32
    assert [sp.distance_to(v) for v in range(6)] == [0, 2, 5, 6, INF, INF]
    check(h, sp)
    h4 = dg.Hypergraph.from_lists(LISTS[:4], WEIGHTS[:4])
    sp4 = hs.compute(h4, 0)
    st = hs.update(h4, dg.HyperedgeBatch(insert_incidences=([0], [3])), sp4)
    assert sp4.hyperedge_distances.tolist() == [2, 5, 3, 9] and st.line_graph_insertions == 1


def test_update_with_hyperedges_and_incidence_edits():
    """test_hyper_sssp.py:104-130."""
    _say("[hyper_sssp] update with hyperedges and incidence edits")
    h = dg.Hypergraph.from_lists(LISTS, WEIGHTS)
    sp = hs.compute(h, 0)
    st = hs.update(h, dg.HyperedgeBatch(insert_incidences=([0], [3])), sp)
    assert sp.hyperedge_distances.tolist()[:4] == [2, 5, 3, 9]
    assert (st.line_graph_insertions, st.line_graph_deletions) == (1, 0)
    assert st.batch.inserted_incidences == 1 and st.batch.skipped_ops == 0
    assert st.engine_used == "operators" and st.converged and not st.fallback_used
    assert st.fallback_reason == "none"
    check(h, sp)
    st = hs.update(h, dg.HyperedgeBatch(delete_hyperedges=[3], insert_hyperedges=[[3, 4]], insert_weights=[1],
                                        delete_incidences=([0, 1], [1, 3])), sp)
    assert st.batch.deleted_incidences == 1 and st.batch.skipped_ops == 1
    assert st.batch.inserted_ids.size == 1 and st.invalidated == st.affected
    check(h, sp)


@pytest.mark.parametrize("reuse", ["erase_first", "insert_first"])
@pytest.mark.parametrize("seed,n,m,kmax,wmax", [(1, 12, 10, 4, 9), (2, 30, 40, 5, 3), (3, 60, 120, 6, 20),
                                               (4, 200, 500, 4, 1000)])
def test_update_chain_equals_compute(reuse, seed, n, m, kmax, wmax):
    """Random weighted hypergraphs (small weights: many equal-distance ties) through batches of
    hyperedge deletions and insertions and incidence edits: after every update the result equals
    the oracle and compute(), and the line graph's pair counts equal the oracle's delta."""
    _say(f"[hyper_sssp chain] reuse={reuse} seed={seed} n={n} m={m}")
    rng = np.random.default_rng(seed)
    lists = [rng.integers(0, n, int(rng.integers(1, kmax + 1))).tolist() for _ in range(m)]
    h = dg.Hypergraph.from_lists(lists, rng.integers(1, wmax + 1, m).tolist(), num_vertices=n, id_reuse=reuse)
    sources = [int(x) for x in rng.integers(0, n, 2)]
    sps = [hs.compute(h, s) for s in sources]
    for sp in sps:
        check(h, sp)
    for step in range(6):
        bound = max(h.hyperedge_id_bound, 1)
        k = max(1, m // 8)
        dels = rng.integers(0, bound, int(rng.integers(0, k + 1))).tolist()
        ins = [rng.integers(0, n, int(rng.integers(1, kmax + 1))).tolist() for _ in range(int(rng.integers(0, k + 1)))]
        w = rng.integers(1, wmax + 1, len(ins)).tolist()
        k1, k2 = int(rng.integers(0, k + 1)), int(rng.integers(0, k + 1))
        rem = (rng.integers(0, bound, k1).tolist(), rng.integers(0, n, k1).tolist())
        add = (rng.integers(0, bound, k2).tolist(), rng.integers(0, n, k2).tolist())
        batch = dg.HyperedgeBatch(dels, ins, w, delete_incidences=rem, insert_incidences=add)
        pre = h.rows()
        # the first result goes through update(); the second is stale afterwards and recomputed
        st = hs.update(h, batch, sps[0])
        deleted, inserted = oh.line_graph_delta(pre, h.rows())
        assert (st.line_graph_insertions, st.line_graph_deletions) == (len(inserted), len(deleted)), step
        check(h, sps[0])
        with pytest.raises(dg.StaleResultError):
            hs.update(h, dg.HyperedgeBatch(), sps[1])
        sps[1] = hs.compute(h, sources[1])
        check(h, sps[1])


def test_requirements_and_errors():
    _say("[hyper_sssp] errors")
    with pytest.raises(dg.NotSupportedError, match="weighted"):
        hs.compute(dg.Hypergraph.from_lists([[0, 1]]), 0)
    h = dg.Hypergraph.from_lists(LISTS, WEIGHTS)
    with pytest.raises(dg.InvalidArgumentError):
        hs.compute(h, 6)
    sp = hs.compute(h, 0)
    before = (h.version, h.rows())
    with pytest.raises(dg.InvalidArgumentError, match="2\\^28"):
        hs.update(h, dg.HyperedgeBatch(insert_hyperedges=[[0, 1]], insert_weights=[2**28 + 1]), sp)
    assert (h.version, h.rows()) == before
    check(h, sp)
    hs.update(h, dg.HyperedgeBatch(delete_hyperedges=[1]), sp)
    assert sp.hyperedge_distances.tolist()[1] == INF
    h.apply(dg.HyperedgeBatch(delete_hyperedges=[0]))  # the hypergraph moves on without the result
    with pytest.raises(dg.StaleResultError):
        sp.distance_to(1)
    with pytest.raises(dg.StaleResultError):
        hs.update(h, dg.HyperedgeBatch(), sp)
    sp2 = hs.compute(h, 0)
    copy = sp2.clone()
    assert copy.hyperedge_distances.tolist() == sp2.hyperedge_distances.tolist()
    check(h, copy)
