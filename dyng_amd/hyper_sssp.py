"""dyng_amd.hyper_sssp: dynamic single-source shortest hyperpaths on the shared sssp engine
(python/dyng/hyper_sssp.py; the reference's "operators" engine, which runs sssp's relax on the
line graph, cpp/src/algorithms/hyper_sssp/operators.cu:350-357,457).

On the line graph of a weighted hypergraph (one node per live hyperedge; two hyperedges are
adjacent when they share a vertex), entering hyperedge h costs its weight w(h). From a source
vertex s, D(h) is the weight of the cheapest walk h_1, ..., h_k = h with s in h_1, so D(h) = w(h)
for every hyperedge that holds s; the parent of h is -1 when it holds s, else its lowest-id
neighbour on a shortest walk. A vertex's distance is the smallest D(h) over its live
hyperedges, and 0 for the source.

How it runs here: each result keeps the line graph on the device as a ``dyng_amd.Graph`` whose
vertex 0 is a virtual source and vertex h + 1 is hyperedge h: an edge (0, h + 1) of weight w(h)
for every h that holds s, and an edge (g + 1, h + 1) of weight w(h) for every ordered pair of
hyperedges that share a vertex (the node weight of the head, as the reference's relax adds it).
The shortest-path tree from vertex 0 is the tree above: D(h) is the distance of h + 1, and the
lowest-id tight parent of h + 1 is 0 exactly for the seeded hyperedges (every other tight
parent g + 1 would need D(g) = 0). ``compute`` solves it with the fused engine; ``update``
applies the batch to the hypergraph (the library's host container, :mod:`dyng_amd.hypergraph`),
which turns the line graph's change into an edge batch (the pairs it lost and gained, and the in-edges of a
reused id whose weight changed) and repairs the tree with ``dyng_amd.sssp.update`` (the device
apply and the persistent kernel). Afterwards the result equals ``compute`` on the new
hypergraph, distances and parents.

Differences from the reference: the work and iteration budgets (``Options.work_budget``,
``max_iterations``, ``on_budget``) are kept but never act (the engine runs every update to the
end: ``fallback_used`` is False); ``invalidated``, ``iterations``, ``frontier_visits`` and
``relaxations`` count the engine's own work on the line graph.
"""

from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass, field, replace
from typing import Any

import numpy as np

from . import _lib
from . import sssp as _sssp
from .batch import EdgeBatch
from .errors import InvalidArgumentError, NotSupportedError, StaleResultError
from .graph import Graph
from .hypergraph import MAX_WEIGHT, HyperApplySummary, HyperedgeBatch, Hypergraph

__all__ = ["INFINITE_DISTANCE", "Options", "Stats", "Result", "compute", "update"]

INFINITE_DISTANCE: int = (2**63 - 1) // 4
"""The distance of a dead or unreachable hyperedge or vertex (C++ infinite_distance)."""


@dataclass
class Options:
    """dyng::hyper_sssp::options: tunables, never the result."""

    work_budget: float = 1.0
    max_iterations: int = 4096
    on_budget: str = "recompute"
    cuda_engine: str = "automatic"
    validate_inputs: bool = True

    def _checked(self) -> "Options":
        if isinstance(self.work_budget, bool) or not isinstance(self.work_budget, (int, float, np.integer,
                                                                                   np.floating)):
            raise TypeError("hyper_sssp.Options.work_budget: expected a number")
        if self.work_budget < 0:
            raise InvalidArgumentError("hyper_sssp.Options.work_budget must be >= 0")
        if int(self.max_iterations) < 1:
            raise InvalidArgumentError("hyper_sssp.Options.max_iterations must be >= 1")
        if self.on_budget not in ("recompute", "error"):
            raise InvalidArgumentError(f"hyper_sssp.Options.on_budget: {self.on_budget!r} is not 'recompute' or 'error'")
        if self.cuda_engine not in ("automatic", "fused", "operators"):
            raise InvalidArgumentError(f"hyper_sssp.Options.cuda_engine: unknown engine {self.cuda_engine!r}")
        return self


@dataclass(frozen=True)
class Stats:
    """dyng::hyper_sssp::stats of one update."""

    affected: int
    iterations: int
    frontier_visits: int
    fallback_used: bool
    converged: bool
    engine_used: str
    batch: HyperApplySummary
    line_graph_insertions: int
    line_graph_deletions: int
    invalidated: int
    relaxations: int
    fallback_reason: str = "none"
    device: Any = field(default=None, compare=False, repr=False)  # the line-graph repair's sssp stats


def _line_graph_edges(hg: Hypergraph, source: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Every edge of a result's device line graph (dyng_hip_hyper_line_graph): (0, h + 1) for the
    hyperedges that hold the source, (g + 1, h + 1) for the ordered pairs that share a vertex;
    weight w(h)."""
    count = C.c_int64()
    src, dst, w = _lib._i32p(), _lib._i32p(), _lib._i32p()
    _lib.check(_lib.lib().dyng_hip_hyper_line_graph(hg._h, int(source), C.byref(count), C.byref(src), C.byref(dst),
                                                    C.byref(w)))
    m = count.value
    if m == 0:
        z = np.zeros(0, np.int32)
        return z, z.copy(), z.copy()
    return tuple(np.ctypeslib.as_array(x, shape=(m,)).copy() for x in (src, dst, w))


def _edge_batch(hg: Hypergraph, source: int) -> tuple[EdgeBatch, int, int]:
    """The change of a result's device line graph over the hypergraph's last apply
    (dyng_hip_hyper_line_graph_delta), and the unordered pairs the line graph gained and lost."""
    d = _lib.HyperDelta()
    _lib.check(_lib.lib().dyng_hip_hyper_line_graph_delta(hg._h, int(source), C.byref(d)))

    def arr(p, k):
        return np.ctypeslib.as_array(p, shape=(k,)).copy() if k else np.zeros(0, np.int32)

    ni, nd = d.num_insertions, d.num_deletions
    batch = EdgeBatch._from_library((arr(d.insert_src, ni), arr(d.insert_dst, ni), arr(d.insert_weights, ni)),
                                    (arr(d.delete_src, nd), arr(d.delete_dst, nd)))
    return batch, int(d.line_graph_insertions), int(d.line_graph_deletions)


class Result:
    """Shortest hyperpaths from one source vertex, kept up to date by :func:`update`
    (dyng::hyper_sssp::result)."""

    __slots__ = ("_source", "_options", "_lg", "_sp", "_hg", "_version", "_bound", "_poisoned", "__weakref__")

    def __init__(self) -> None:
        raise TypeError("use dyng_amd.hyper_sssp.compute()")

    @property
    def source(self) -> int:
        return self._source

    def _arrays(self) -> tuple[np.ndarray, np.ndarray]:
        if self._poisoned:
            raise StaleResultError("dyng_amd.hyper_sssp.Result: a failed update left this result unusable; "
                                   "recompute it")
        d, p = self._sp.to_numpy()
        b = self._bound
        dist = np.full(b, INFINITE_DISTANCE, np.int64)
        par = np.full(b, -1, np.int32)
        k = min(b, len(d) - 1)
        if k > 0:
            dist[:k] = d[1:k + 1]
            pp = p[1:k + 1].astype(np.int64)
            par[:k] = np.where(pp > 0, pp - 1, -1).astype(np.int32)
        dist[dist >= INFINITE_DISTANCE] = INFINITE_DISTANCE
        dist.flags.writeable = False
        par.flags.writeable = False
        return dist, par

    @property
    def hyperedge_distances(self) -> np.ndarray:
        """D(h) by hyperedge id (int64, hyperedge_id_bound values); INF for dead or unreachable ids."""
        return self._arrays()[0]

    @property
    def hyperedge_parents(self) -> np.ndarray:
        """The canonical parent of every id (int32); -1 for a hyperedge reached from the source
        itself and for dead or unreachable ids."""
        return self._arrays()[1]

    def _graph(self) -> Hypergraph:
        hg = self._hg()
        if hg is None or hg.version != self._version:
            raise StaleResultError("dyng_amd.hyper_sssp.Result: the hypergraph this result matches is gone or "
                                   "has changed since (update the result with the batch)")
        return hg

    def distance_to(self, target: int) -> int:
        """The smallest D(h) over the live hyperedges of `target`, 0 for the source."""
        v = int(target)
        hg = self._graph()
        if not 0 <= v < hg.num_vertices:
            raise InvalidArgumentError(f"hyper_sssp.Result.distance_to: the vertex {v} is outside [0, num_vertices)")
        if v == self._source:
            return 0
        d = self._arrays()[0]
        return int(min((d[h] for h, row in enumerate(hg.rows()) if v in row), default=INFINITE_DISTANCE))

    def vertex_distances(self) -> np.ndarray:
        """The distance of every vertex (int64, host memory)."""
        hg = self._graph()
        d = self._arrays()[0]
        out = np.full(hg.num_vertices, INFINITE_DISTANCE, np.int64)
        for h, row in enumerate(hg.rows()):
            if row:
                idx = np.asarray(row, np.int64)
                out[idx] = np.minimum(out[idx], d[h])
        if 0 <= self._source < hg.num_vertices:
            out[self._source] = 0
        out.flags.writeable = False
        return out

    @property
    def options(self) -> Options:
        return replace(self._options)

    def set_options(self, options: Options | None = None, **kwargs: Any) -> None:
        self._options = replace(options or self._options, **kwargs)._checked()

    @property
    def graph_version(self) -> int:
        return self._version

    @property
    def space(self) -> str:
        return "device"

    def clone(self) -> "Result":
        """A copy with a line graph of its own (built again from the hypergraph, which it must
        still match: the copy equals compute() on it, as this result does)."""
        if self._poisoned:
            raise StaleResultError("dyng_amd.hyper_sssp.Result.clone: a failed update left this result unusable")
        return compute(self._graph(), self._source, options=self._options)

    def __reduce__(self) -> Any:
        raise TypeError("dyng_amd.hyper_sssp.Result cannot be pickled; send hyperedge_distances and "
                        "hyperedge_parents")

    def __repr__(self) -> str:
        return (f"dyng_amd.hyper_sssp.Result(source={self._source}, num_hyperedge_ids={self._bound}, "
                f"graph_version={self._version})")


def _check_hypergraph(hg: Any) -> Hypergraph:
    if not isinstance(hg, Hypergraph):
        raise TypeError("dyng_amd.hyper_sssp: expected a dyng_amd.Hypergraph")
    if not hg.weighted:
        raise NotSupportedError("dyng_amd.hyper_sssp needs a weighted hypergraph (int32 weights in [1, 2^28], the "
                                "hyperedge lengths); this hypergraph is unweighted (build it with weights)")
    return hg


def compute(hypergraph: Hypergraph, source: int, *, options: Options | None = None, resources: Any = None,
            **kwargs: Any) -> Result:
    """Shortest hyperpaths of `hypergraph` from the source vertex `source` (hyper_sssp::compute).
    The hypergraph is not modified."""
    hg = _check_hypergraph(hypergraph)
    opt = replace(options or Options(), **kwargs)._checked()
    s = int(source)
    if not 0 <= s < hg.num_vertices:
        raise InvalidArgumentError(f"hyper_sssp::compute: source {s} is out of range [0, {hg.num_vertices})")
    w = hg.weights
    if opt.validate_inputs and w is not None:
        live = np.asarray([bool(r) for r in hg.rows()], bool)
        if live.any() and (int(w[live].min()) < 1 or int(w[live].max()) > MAX_WEIGHT):
            raise InvalidArgumentError("hyper_sssp::compute: a live weight is outside [1, 2^28]")
    bound = hg.hyperedge_id_bound
    src, dst, wt = _line_graph_edges(hg, s)
    lg = Graph.from_edges(src, dst, wt, num_vertices=bound + 1)
    r = object.__new__(Result)
    r._source = s
    r._options = opt
    r._lg = lg
    r._sp = _sssp.compute(lg, 0)
    r._hg = weakref.ref(hg)
    r._version = hg.version
    r._bound = bound
    r._poisoned = False
    return r


def update(hypergraph: Hypergraph, batch: HyperedgeBatch, result: Result, *, resources: Any = None) -> Stats:
    """Apply `batch` to `hypergraph` and update `result` (hyper_sssp::update). Afterwards the result
    equals compute() on the new hypergraph."""
    hg = _check_hypergraph(hypergraph)
    if not isinstance(result, Result):
        raise TypeError("dyng_amd.hyper_sssp.update: result must be a dyng_amd.hyper_sssp.Result")
    if not isinstance(batch, HyperedgeBatch):
        raise TypeError("dyng_amd.hyper_sssp.update: batch must be a dyng_amd.HyperedgeBatch")
    if result._poisoned or result._hg() is not hg or result._version != hg.version:
        raise StaleResultError("dyng_amd.hyper_sssp.update: the result does not match the hypergraph's current "
                               "state (it was computed on another hypergraph, or the hypergraph changed without it)")
    summary = hg._apply(batch, MAX_WEIGHT if result._options.validate_inputs else 0)
    try:
        eb, gained, lost = _edge_batch(hg, result._source)
        st = _sssp.update(result._lg, eb, result._sp)
    except BaseException:
        result._poisoned = True
        raise
    result._version = hg.version
    result._bound = hg.hyperedge_id_bound
    dev = st.device
    return Stats(affected=st.invalidated, iterations=st.iterations, frontier_visits=st.frontier_visits,
                 fallback_used=False, converged=True, engine_used="operators", batch=summary,
                 line_graph_insertions=gained, line_graph_deletions=lost, invalidated=st.invalidated,
                 relaxations=int(dev.relax_edges + dev.pull_edges), device=st)
