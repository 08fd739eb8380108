"""dyng_amd.mosp: multi-objective shortest paths (the MOSP update of DynaMOSP) on the HIP path.

A graph with K weight columns (objectives) and a source. :func:`compute` builds the canonical
shortest-path tree of every objective k (:func:`dyng_amd.sssp.compute` with ``objective=k``),
then the *combined graph* of the K trees, whose edge (p, v) weighs ``L (K + 1) - sum L / Pref_i``
over the trees i that hold it (``Pref`` the preferences, a lower value a higher priority,
``L = lcm(Pref)``), its canonical shortest-path tree (the **MOSP tree**) and the **path costs**,
the K objective values of the MOSP path to every vertex. :func:`update` applies a batch once,
repairs the K trees on the device and rebuilds the combined graph, the MOSP tree and the costs.

Names and behaviour follow the reference's python/dyng/mosp.py (Options :93-136, Stats :139-183,
Result :186-403, compute :420-455, update :458-509) over the C++ path of
cpp/src/algorithms/mosp/mosp.cpp:264-330: the combined graph and the path costs are built by the
kernels of mosp.hip, the combined tree by the persistent sssp engine. A result keeps its combined
graph on the device (``dyng_hip_mosp_create``): an update turns it into the new trees' combined
graph with one batch and repairs its tree (``dyng_hip_mosp_update``), the same tree the
reference's solve from scratch gives.
The trees and the combined arrays stay in device memory; the path costs are returned as host
memory, as in the reference.

Worked example (thesis Chapter 4, vertices u1..u7 are 0..6; python/dyng/mosp.py:22-37): with
Pref (4, 1, 4) the MOSP path to u7 costs (15, 3, 20) after the batch.
"""

from __future__ import annotations

import ctypes as C
import dataclasses
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from . import _lib, sssp as _sssp
from .batch import EdgeBatch
from .device import DeviceArray
from .errors import InvalidArgumentError, NotSupportedError, StaleResultError
from .graph import ApplySummary, Graph
from .resources import Resources, resolve

__all__ = ["Options", "Stats", "Result", "compute", "update", "INFINITE_DISTANCE", "MAX_OBJECTIVES",
           "MAX_PREFERENCE_SCALE"]

INFINITE_DISTANCE = _sssp.INFINITE_DISTANCE
MAX_OBJECTIVES = 64
MAX_PREFERENCE_SCALE = 1 << 20
_INT32_MAX = 2**31 - 1


def _preferences(value: Any, what: str) -> list[int]:
    if isinstance(value, (str, bytes)):
        raise InvalidArgumentError(f"{what}: expected a sequence of integers, got {value!r}")
    try:
        out = [int(p) for p in value]
    except TypeError:
        raise InvalidArgumentError(f"{what}: expected a sequence of integers, got {value!r}") from None
    for p in out:
        if not 1 <= p <= _INT32_MAX:
            raise InvalidArgumentError(f"{what}: preferences must be in [1, 2^31 - 1], got {p}")
    return out


@dataclass
class Options:
    """mosp::options (python/dyng/mosp.py:93-136).

    preferences: one per objective, each >= 1 (lower = higher priority); empty = all ones; lcm at
    most 2^20; fixed at compute. delta: bucket width of the K sssp updates (0 = automatic per
    objective; the combined solve always takes the automatic width of the combined graph).
    cuda_engine: the engine of the K updates. compute_path_costs: whether compute / update fill
    :attr:`Result.path_costs`. validate_inputs: the checks of trees imported with
    :meth:`Result.from_arrays`. num_objectives: the objectives are the first num_objectives weight
    columns; 0 = every column; fixed at compute.
    """

    preferences: list[int] = field(default_factory=list)
    delta: int = 0
    cuda_engine: str = "automatic"
    compute_path_costs: bool = True
    validate_inputs: bool = True
    num_objectives: int = 0

    def _checked(self, graph: Graph, what: str) -> "Options":
        prefs = _preferences(self.preferences, f"{what}: preferences")
        k = int(self.num_objectives)
        if k < 0 or k > graph.num_weights:
            raise InvalidArgumentError(f"{what}: num_objectives {k} is out of range [0, {graph.num_weights}]")
        count = k or graph.num_weights
        if count > MAX_OBJECTIVES:
            raise InvalidArgumentError(f"{what}: {count} objectives; at most {MAX_OBJECTIVES}")
        if prefs and len(prefs) != count:
            raise InvalidArgumentError(f"{what}: {len(prefs)} preferences for {count} objectives")
        if int(self.delta) < 0:
            raise InvalidArgumentError(f"{what}: delta must be >= 0")
        if self.cuda_engine not in _lib.ENGINES:
            raise InvalidArgumentError(f"{what}: unknown engine {self.cuda_engine!r}")
        scale = 1
        for p in prefs:
            a, b = scale, p
            while b:
                a, b = b, a % b
            scale = scale // a * p
            if scale > MAX_PREFERENCE_SCALE:
                raise InvalidArgumentError(f"{what}: lcm(preferences) exceeds 2^20")
        return dataclasses.replace(self, preferences=prefs, num_objectives=k)

    def _sssp(self, objective: int) -> _sssp.Options:
        return _sssp.Options(delta=int(self.delta), objective=objective, cuda_engine=self.cuda_engine,
                             validate_inputs=bool(self.validate_inputs))


@dataclass(frozen=True)
class Stats:
    """mosp::stats (python/dyng/mosp.py:139-183): affected counts the vertices whose combined
    distance or MOSP parent changed; iterations and frontier_visits are summed over the K sssp
    updates and the combined solve; engine_used is the first objective's."""

    affected: int
    iterations: int
    frontier_visits: int
    fallback_used: bool
    converged: bool
    engine_used: str
    batch: ApplySummary
    objectives: tuple
    combined_edges: int
    preference_scale: int
    combined_solve: Any = None  # dyng_amd.sssp.DeviceCounters of the combined tree's repair (HIP path only)
    combined_batch: tuple = (0, 0)  # the combined graph's batch: (deletions, insertions or re-weights)


class Result:
    """The K trees, the MOSP tree and the path costs kept up to date by :func:`update`
    (mosp::result, python/dyng/mosp.py:186-403)."""

    __slots__ = ("_trees", "_options", "_resources", "_source", "_count", "_state", "_graph", "_costs",
                 "_costs_host", "_costs_valid", "_scale", "_combined_edges", "__weakref__")

    def __init__(self):
        raise TypeError("use dyng_amd.mosp.compute() or dyng_amd.mosp.Result.from_arrays()")

    @classmethod
    def _new(cls, trees: list, options: Options, count: int, source: int, res: Resources) -> "Result":
        self = object.__new__(cls)
        self._trees = trees
        self._options = options
        self._resources = res
        self._source = int(source)
        self._count = count
        self._state = None  # dyng_hip_mosp*: the combined graph and its tree on the device
        self._graph = None  # the graph the result belongs to
        self._costs = None  # device (n, K) int64
        self._costs_host = None  # its host copy, made on first read after each compute / update
        self._costs_valid = False
        self._scale = 1
        self._combined_edges = 0
        return self

    def _objective(self, objective: int, what: str) -> int:
        k = int(objective)
        if not 0 <= k < self._count:
            raise InvalidArgumentError(f"mosp.Result.{what}: the objective {k} is out of range [0, {self._count})")
        return k

    @property
    def source(self) -> int:
        return self._source

    @property
    def num_objectives(self) -> int:
        return self._count

    def tree(self, objective: int) -> _sssp.Result:
        """Objective k's shortest-path tree (a device-resident sssp result)."""
        return self._trees[self._objective(objective, "tree")]

    def distances(self, objective: int) -> np.ndarray:
        """Host copy of objective k's distances (INFINITE_DISTANCE where unreachable)."""
        return self.tree(objective).distances

    def parents(self, objective: int) -> np.ndarray:
        """Host copy of objective k's parents (-1 for the source and unreachable vertices)."""
        return self.tree(objective).parents

    def combined_device_arrays(self) -> tuple[int, int, int]:
        """(combined distances pointer, MOSP parents pointer, n) in device memory."""
        d, p, n = C.c_void_p(), C.c_void_p(), C.c_int64()
        if self._state is None:
            return 0, 0, 0
        _lib.check(_lib.lib().dyng_hip_mosp_arrays(self._state, C.byref(d), C.byref(p), C.byref(n)))
        return d.value or 0, p.value or 0, n.value

    def _download(self, which: int, dtype) -> np.ndarray:
        ptrs = self.combined_device_arrays()
        out = np.empty(ptrs[2], dtype)
        if out.nbytes:
            _lib.check(_lib.lib().dyng_hip_memcpy_dtoh(out.ctypes.data, ptrs[which], out.nbytes))
        return out

    @property
    def combined_distances(self) -> np.ndarray:
        """Host copy of the combined distances (units of 1/L, L = preference_scale)."""
        return self._download(0, np.int64)

    @property
    def combined_parents(self) -> np.ndarray:
        """Host copy of the MOSP tree (-1 for the source and unreachable vertices)."""
        return self._download(1, np.int32)

    @property
    def path_costs(self) -> np.ndarray:
        """(n, K) int64 host array: the K objective values of every vertex's MOSP path
        (INFINITE_DISTANCE where the MOSP tree does not reach). Computed on the device by every
        compute / update; copied to the host on the first read after it."""
        if not self._costs_valid:
            raise InvalidArgumentError("mosp.Result.path_costs: the last compute or update ran with "
                                       "compute_path_costs=False")
        if self._costs_host is None:
            self._costs_host = self._costs.to_numpy() if self._costs is not None else \
                np.zeros((0, self._count), np.int64)
        return self._costs_host.copy()

    def path_costs_device(self) -> tuple[int, int, int]:
        """(path costs pointer, n, K) in device memory (vertex-major, int64)."""
        if not self._costs_valid:
            raise InvalidArgumentError("mosp.Result.path_costs_device: the last compute or update ran with "
                                       "compute_path_costs=False")
        n = self._costs.shape[0] if self._costs is not None else 0
        return (self._costs.ptr if self._costs is not None else 0, n, self._count)

    @property
    def preference_scale(self) -> int:
        return self._scale

    @property
    def combined_edges(self) -> int:
        return self._combined_edges

    @property
    def options(self) -> Options:
        return dataclasses.replace(self._options, preferences=list(self._options.preferences))

    def set_options(self, options: Options | None = None, **kwargs: Any) -> None:
        """Change delta, cuda_engine, compute_path_costs and validate_inputs for the next update;
        preferences and num_objectives stay fixed."""
        opt = dataclasses.replace(options or self.options, **kwargs)
        if list(opt.preferences) != list(self._options.preferences) or \
                int(opt.num_objectives) != int(self._options.num_objectives):
            raise InvalidArgumentError("mosp.Result.set_options: preferences and num_objectives are fixed at compute")
        if int(opt.delta) < 0:
            raise InvalidArgumentError("mosp.Result.set_options: delta must be >= 0")
        if opt.cuda_engine not in _lib.ENGINES:
            raise InvalidArgumentError(f"mosp.Result.set_options: unknown engine {opt.cuda_engine!r}")
        for k, t in enumerate(self._trees):
            t.set_options(opt._sssp(k))
        self._options = opt

    @property
    def graph_version(self) -> int:
        return self._trees[0].graph_version

    @property
    def space(self) -> str:
        return "device"

    @property
    def vertex_dtype(self) -> np.dtype:
        return np.dtype(np.int32)

    def clone(self, resources: Resources | None = None) -> "Result":
        """A deep copy (trees, combined graph and tree, costs, options) of a result that matches
        its graph's current version."""
        if self._graph is None or self.graph_version != self._graph.version:
            raise StaleResultError("dyng_amd.mosp.Result.clone: the result does not match its graph's current "
                                   "version")
        res = resolve(resources, self._resources)
        out = Result._new([t.clone(res) for t in self._trees], self.options, self._count, self._source, res)
        _combine(self._graph, out, update=False)
        return out

    def __del__(self):
        h = getattr(self, "_state", None)
        if h and _lib._lib is not None:
            _lib._lib.dyng_hip_mosp_destroy(h)
            self._state = None

    @classmethod
    def from_arrays(cls, graph: Graph, source: int, distances, parents, *, canonicalize: bool = True,
                    options: Options | None = None, resources: Resources | None = None,
                    **kwargs: Any) -> "Result":
        """Adopt K existing trees of ``graph`` (python/dyng/mosp.py:319-382) and build the combined
        graph, the MOSP tree and the path costs."""
        _check_graph(graph)
        opt = dataclasses.replace(options or Options(), **kwargs)._checked(graph, "mosp.Result.from_arrays")
        res = resolve(resources, graph.resources)
        if isinstance(distances, (str, bytes)) or isinstance(parents, (str, bytes)):
            raise TypeError("mosp.Result.from_arrays: distances and parents are lists of arrays")
        distances, parents = list(distances), list(parents)
        count = opt.num_objectives or graph.num_weights
        if len(distances) != count or len(parents) != count:
            raise InvalidArgumentError(f"mosp.Result.from_arrays: {len(distances)} distance and {len(parents)} "
                                       f"parent arrays for {count} objectives")
        trees = [_sssp.Result.from_arrays(graph, source, distances[k], parents[k], canonicalize=canonicalize,
                                          options=opt._sssp(k), resources=res) for k in range(count)]
        out = cls._new(trees, opt, count, source, res)
        _combine(graph, out, update=False)
        return out

    def __copy__(self) -> "Result":
        return self.clone()

    def __deepcopy__(self, memo) -> "Result":
        return self.clone()

    def __reduce__(self):
        raise TypeError("dyng_amd.mosp.Result cannot be pickled; send the K trees and rebuild it with "
                        "dyng_amd.mosp.Result.from_arrays()")

    def __repr__(self) -> str:
        n = self.combined_device_arrays()[2]
        return (f"dyng_amd.mosp.Result(source={self._source}, num_objectives={self._count}, num_vertices={n}, "
                f"preference_scale={self._scale}, graph_version={self.graph_version})")


def _check_graph(graph: Graph) -> None:
    if not isinstance(graph, Graph):
        raise TypeError("expected a dyng_amd.Graph")
    if graph.num_weights < 1:
        raise NotSupportedError("dyng_amd.mosp needs integer weights: one int32 weight column per objective")


def _combine(graph: Graph, result: Result, *, update: bool) -> _lib.MospStats:
    """The combined graph, the MOSP tree and the path costs of ``result``'s K trees
    (mosp.cpp:264-330): built on the device (compute, from_arrays), or, after the trees were
    repaired, brought up to date with one batch and a repair (update)."""
    res = result._resources
    n = graph.num_vertices
    count = result._count
    ptrs = (C.c_void_p * count)(*[t.device_arrays()[1] for t in result._trees])
    st = _lib.MospStats()
    if update and result._state is not None:
        _lib.check(_lib.lib().dyng_hip_mosp_update(result._state, graph._handle, ptrs, res.stream_handle,
                                                   C.byref(st)))
    else:
        prefs = result._options.preferences
        pref_arr = (C.c_int32 * count)(*prefs) if prefs else None
        h = C.c_void_p()
        _lib.check(_lib.lib().dyng_hip_mosp_create(graph._handle, result._source, count, ptrs, pref_arr,
                                                   res.stream_handle, C.byref(h), C.byref(st)))
        if result._state is not None:
            _lib.lib().dyng_hip_mosp_destroy(result._state)
        result._state = h.value
    result._graph = graph
    columns = count if result._options.compute_path_costs else 0
    costs = result._costs
    if columns and n and (costs is None or costs.shape != (n, columns)):
        costs = DeviceArray.empty((n, columns), np.int64, device=res.device)
    if columns and n:
        _lib.check(_lib.lib().dyng_hip_mosp_path_costs(result._state, graph._handle, columns, costs.ptr,
                                                       res.stream_handle))
    result._costs_host = None
    if columns:
        result._costs = costs if n else None
        result._costs_valid = True
    else:
        result._costs_valid = False
    result._scale = int(st.preference_scale)
    result._combined_edges = int(st.combined_edges)
    return st


def path_costs_over(graph: Graph, result: Result, columns: int) -> np.ndarray:
    """The first ``columns`` weight columns of ``graph`` (up to all of them, more than the
    objectives) summed along ``result``'s MOSP tree: (n, columns) int64 on the host, as MOSP's
    mospCosts.txt holds them. ``result`` must match the graph's current version."""
    if not isinstance(result, Result):
        raise TypeError("dyng_amd.mosp.path_costs_over: result must be a dyng_amd.mosp.Result")
    if result.graph_version != graph.version:
        raise StaleResultError("dyng_amd.mosp.path_costs_over: the result does not match the graph's version")
    columns = int(columns)
    if not 1 <= columns <= graph.num_weights:
        raise InvalidArgumentError(f"mosp.path_costs_over: columns must be in [1, {graph.num_weights}]")
    res = result._resources
    n = graph.num_vertices
    if n == 0:
        return np.zeros((0, columns), np.int64)
    count = result._count
    ptrs = (C.c_void_p * count)(*[t.device_arrays()[1] for t in result._trees])
    prefs = result._options.preferences
    pref_arr = (C.c_int32 * count)(*prefs) if prefs else None
    dist = DeviceArray.empty(n, np.int64, device=res.device)
    par = DeviceArray.empty(n, np.int32, device=res.device)
    costs = DeviceArray.empty((n, columns), np.int64, device=res.device)
    st = _lib.MospStats()
    _lib.check(_lib.lib().dyng_hip_mosp_combined(graph._handle, result._source, count, ptrs, pref_arr, 0,
                                                 dist.ptr, par.ptr, columns, costs.ptr, res.stream_handle,
                                                 C.byref(st)))
    return costs.to_numpy()


def compute(graph: Graph, source: int, *, options: Options | None = None, resources: Resources | None = None,
            **kwargs: Any) -> Result:
    """The K canonical trees of ``graph`` from ``source``, the combined graph, the MOSP tree and
    the path costs (mosp::compute, python/dyng/mosp.py:420-455)."""
    _check_graph(graph)
    for key in kwargs:
        if key not in {f.name for f in dataclasses.fields(Options)}:
            raise TypeError(f"mosp.compute: unexpected keyword argument {key!r}")
    opt = dataclasses.replace(options or Options(), **kwargs)._checked(graph, "mosp.compute")
    res = resolve(resources, graph.resources)
    count = opt.num_objectives or graph.num_weights
    trees = [_sssp.compute(graph, source, options=opt._sssp(k), resources=res) for k in range(count)]
    out = Result._new(trees, opt, count, source, res)
    _combine(graph, out, update=False)
    return out


def _finish_update(graph: Graph, result: Result, per: tuple) -> Stats:
    """After the K trees were repaired on the new graph: the combined graph, the MOSP tree and
    the costs, and the update's stats."""
    st = _combine(graph, result, update=True)
    solve = st.solve
    return Stats(affected=int(st.affected),
                 iterations=sum(s.iterations for s in per) + int(solve.iterations),
                 frontier_visits=sum(s.frontier_visits for s in per) + int(solve.pushes),
                 fallback_used=any(s.fallback_used for s in per),
                 converged=all(s.converged for s in per),
                 engine_used=per[0].engine_used if per else "fused",
                 batch=per[0].batch, objectives=per, combined_edges=int(st.combined_edges),
                 preference_scale=int(st.preference_scale),
                 combined_solve=_sssp.DeviceCounters._from_native(solve),
                 combined_batch=(int(st.deletions), int(st.insertions)))


def update(graph: Graph, batch: EdgeBatch, result: Result, *, resources: Resources | None = None) -> Stats:
    """Apply ``batch`` to ``graph`` once, repair the K trees and rebuild the combined graph, the
    MOSP tree and the path costs (mosp::update, python/dyng/mosp.py:458-509). Afterwards
    ``result`` equals ``compute(graph, result.source, options=result.options)``. The batch is
    applied and the trees repaired in one call (dyng::update_each); a stale result changes
    nothing."""
    from ._update import update_trees

    _check_graph(graph)
    if not isinstance(result, Result):
        raise TypeError("dyng_amd.mosp.update: result must be a dyng_amd.mosp.Result")
    if not isinstance(batch, EdgeBatch):
        raise TypeError("dyng_amd.mosp.update: batch must be a dyng_amd.EdgeBatch")
    per = tuple(update_trees(graph, batch, result._trees, resources))
    return _finish_update(graph, result, per)
