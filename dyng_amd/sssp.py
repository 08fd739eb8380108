"""dyng_amd.sssp: batch-dynamic single-source shortest paths on MI355X (python/dyng/sssp.py).

Mirrors the reference's Python layer: :class:`Options`, :class:`Result`, :class:`Stats`,
:func:`compute`, :func:`update` and ``Result.from_arrays`` (python/dyng/sssp.py:50-386, bound by
python/bindings/sssp.cpp:77-135). The work runs in libdyng_hip.so (HIP, gfx950): the batch is
applied to the device graph in place and the tree is repaired by the persistent engine. Distances
and parents equal the reference's on every input (the tie rule of cpp/include/dyng/sssp.hpp:34-47:
canonical trees stay canonical; the device packing rule of the cuda backend, ADR 0029).
"""

from __future__ import annotations

import ctypes as C
import dataclasses
from dataclasses import dataclass
from typing import Any

import numpy as np

from . import _lib
from .batch import EdgeBatch
from .errors import InvalidArgumentError, StaleResultError
from .graph import ApplySummary, Graph
from .resources import Resources, resolve

__all__ = ["Options", "Stats", "Result", "Tuning", "compute", "update", "INFINITE_DISTANCE"]

INFINITE_DISTANCE: int = int(np.iinfo(np.int64).max) // 4
"""The "unreachable" distance, max(int64) // 4 = 2**61 - 1 (MOSP's DISTANCE_INF)."""


@dataclass
class Options:
    """sssp::options (cpp/include/dyng/sssp.hpp:78-94)."""

    delta: int = 0
    objective: int = 0
    cuda_engine: str = "automatic"
    validate_inputs: bool = True

    def _native(self) -> _lib.Options:
        if self.cuda_engine not in _lib.ENGINES:
            raise InvalidArgumentError(f"sssp.Options.cuda_engine: unknown engine {self.cuda_engine!r}")
        return _lib.Options(int(self.delta), int(self.objective), _lib.ENGINES[self.cuda_engine],
                            int(bool(self.validate_inputs)))

    @classmethod
    def _from_native(cls, o: _lib.Options) -> "Options":
        return cls(int(o.delta), int(o.objective), _lib.ENGINE_NAMES.get(o.cuda_engine, "automatic"),
                   bool(o.validate_inputs))


@dataclass
class Tuning:
    """Schedule knobs of the HIP engine (they never change the result)."""

    engine: str = "automatic"
    blocks_per_cu: int = 0
    light_degree: int = 0
    chunk_edges: int = 0
    launch: str = "plain"  # "plain" or "graph" (the launch built once as a hipGraph and replayed)
    recompute: bool = False  # updates recompute the tree from scratch instead of repairing it
    barrier_timeout_us: int = 0
    grid_blocks: int = 0  # workgroups of the persistent launch (0: blocks_per_cu per CU)

    def _native(self) -> _lib.Tuning:
        launches = {"plain": 0, "graph": 1}
        if self.launch not in launches:
            raise InvalidArgumentError(f"Tuning.launch must be 'plain' or 'graph', got {self.launch!r}")
        return _lib.Tuning(_lib.ENGINES[self.engine], int(self.grid_blocks), int(self.blocks_per_cu),
                           int(self.light_degree),
                           int(self.chunk_edges), launches[self.launch], int(bool(self.recompute)),
                           int(self.barrier_timeout_us))


@dataclass(frozen=True)
class DeviceCounters:
    invalidated: int
    iterations: int
    epochs: int
    pushes: int
    affected: int
    packed_parents: bool
    relax_vertices: int
    relax_edges: int
    relax_improved: int
    pull_edges: int
    jump_rounds: int
    jump_items: int
    pull_vertices: int
    step1_ms: float
    pull_ms: float
    relax_ms: float
    blocks: int
    kernel_ms: float
    pack_ms: float
    jump_ms: float
    unpack_ms: float
    words_kept: bool = False  # the workspace kept the tree's packed words from the last run (no pack)

    @classmethod
    def _from_native(cls, d) -> "DeviceCounters":
        return cls(int(d.invalidated), int(d.iterations), int(d.epochs), int(d.pushes), int(d.affected),
                   bool(d.packed_parents), int(d.relax_vertices), int(d.relax_edges), int(d.relax_improved),
                   int(d.pull_edges), int(d.jump_rounds), int(d.jump_items), int(d.pull_vertices),
                   float(d.step1_ms), float(d.pull_ms), float(d.relax_ms), int(d.blocks), float(d.kernel_ms),
                   float(d.pack_ms), float(d.jump_ms), float(d.unpack_ms), bool(d.words_kept))


class Stats:
    """sssp::stats (sssp.hpp:100-117) plus the device counters and stream times of this update.

    Read-only. The fields decode from the library's record when first read (an update's stats
    cost the caller no Python work it does not use): ``affected``, ``iterations``,
    ``frontier_visits``, ``fallback_used``, ``converged``, ``engine_used``, ``batch``
    (:class:`dyng_amd.ApplySummary`), ``invalidated``, ``epochs``, ``pushes``,
    ``packed_parents``, ``device`` (:class:`DeviceCounters`), ``apply_ms``, ``engine_ms``.
    """

    __slots__ = ("_raw", "_batch", "_device")
    FIELDS = ("affected", "iterations", "frontier_visits", "fallback_used", "converged", "engine_used", "batch",
              "invalidated", "epochs", "pushes", "packed_parents", "device", "apply_ms", "engine_ms")

    def __init__(self):
        raise TypeError("dyng_amd.sssp.Stats comes from update, repair and dyng_amd.update")

    @classmethod
    def _from_native(cls, s: _lib.Stats) -> "Stats":
        self = object.__new__(cls)
        self._raw = s
        self._batch = None
        self._device = None
        return self

    affected = property(lambda self: int(self._raw.affected))
    iterations = property(lambda self: int(self._raw.iterations))
    frontier_visits = property(lambda self: int(self._raw.frontier_visits))
    fallback_used = property(lambda self: bool(self._raw.fallback_used))
    converged = property(lambda self: bool(self._raw.converged))
    engine_used = property(lambda self: _lib.ENGINE_NAMES.get(self._raw.engine_used, "fused"))
    invalidated = property(lambda self: int(self._raw.invalidated))
    epochs = property(lambda self: int(self._raw.epochs))
    pushes = property(lambda self: int(self._raw.pushes))
    packed_parents = property(lambda self: bool(self._raw.packed_parents))
    apply_ms = property(lambda self: float(self._raw.apply_ms))
    engine_ms = property(lambda self: float(self._raw.engine_ms))

    @property
    def batch(self) -> ApplySummary:
        if self._batch is None:
            self._batch = ApplySummary._from_native(self._raw.batch)
        return self._batch

    @property
    def device(self) -> DeviceCounters:
        if self._device is None:
            self._device = DeviceCounters._from_native(self._raw.device)
        return self._device

    def __setattr__(self, name, value):
        if name in Stats.__slots__:
            object.__setattr__(self, name, value)
        else:
            raise AttributeError("dyng_amd.sssp.Stats is read-only")

    def __eq__(self, other) -> bool:
        if not isinstance(other, Stats):
            return NotImplemented
        return all(getattr(self, f) == getattr(other, f) for f in Stats.FIELDS)

    __hash__ = None

    def __repr__(self) -> str:
        inner = ", ".join(f"{f}={getattr(self, f)!r}" for f in Stats.FIELDS if f not in ("batch", "device"))
        return f"dyng_amd.sssp.Stats({inner})"


class Result:
    """A shortest-path tree kept up to date by :func:`update` (sssp::result, sssp.hpp:153-310).

    The arrays live in device memory; :attr:`distances` / :attr:`parents` return host copies.
    """

    __slots__ = ("_handle", "_resources", "__weakref__")

    def __init__(self):
        raise TypeError("use dyng_amd.sssp.compute() or dyng_amd.sssp.Result.from_arrays()")

    @classmethod
    def _wrap(cls, handle: C.c_void_p, res: Resources) -> "Result":
        self = object.__new__(cls)
        self._handle = handle
        self._resources = res
        return self

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h and _lib._lib is not None:
            _lib._lib.dyng_hip_result_destroy(h)
            self._handle = None

    @property
    def source(self) -> int:
        s = C.c_int32()
        _lib.check(_lib.lib().dyng_hip_result_source(self._handle, C.byref(s)))
        return s.value

    @property
    def counters(self) -> DeviceCounters:
        """The work counters and timings of the engine run that left this tree (compute, update
        or repair)."""
        c = _lib.Counters()
        _lib.check(_lib.lib().dyng_hip_result_counters(self._handle, C.byref(c)))
        return DeviceCounters._from_native(c)

    def device_arrays(self) -> tuple[int, int, int]:
        """(distances pointer, parents pointer, n) in device memory."""
        d, p, n = C.c_void_p(), C.c_void_p(), C.c_int64()
        _lib.check(_lib.lib().dyng_hip_result_arrays(self._handle, C.byref(d), C.byref(p), C.byref(n)))
        return d.value or 0, p.value or 0, n.value

    def to_numpy(self) -> tuple[np.ndarray, np.ndarray]:
        """Host copies of (distances int64, parents int32)."""
        _, _, n = self.device_arrays()
        d = np.empty(n, np.int64)
        p = np.empty(n, np.int32)
        _lib.check(_lib.lib().dyng_hip_result_download(self._handle, d.ctypes.data, p.ctypes.data))
        return d, p

    @property
    def distances(self) -> np.ndarray:
        return self.to_numpy()[0]

    @property
    def parents(self) -> np.ndarray:
        return self.to_numpy()[1]

    @property
    def options(self) -> Options:
        o = _lib.Options()
        _lib.check(_lib.lib().dyng_hip_result_options(self._handle, C.byref(o)))
        return Options._from_native(o)

    def set_options(self, options: Options | None = None, **kwargs: Any) -> None:
        opt = dataclasses.replace(options or self.options, **kwargs)
        o = opt._native()
        _lib.check(_lib.lib().dyng_hip_result_set_options(self._handle, C.byref(o)))

    def set_tuning(self, tuning: Tuning) -> None:
        t = tuning._native()
        _lib.check(_lib.lib().dyng_hip_result_set_tuning(self._handle, C.byref(t)))

    @property
    def graph_version(self) -> int:
        v = C.c_uint64()
        _lib.check(_lib.lib().dyng_hip_result_version(self._handle, C.byref(v)))
        return v.value

    @property
    def space(self) -> str:
        return "device"

    @property
    def vertex_dtype(self) -> np.dtype:
        return np.dtype(np.int32)

    def clone(self, resources: Resources | None = None) -> "Result":
        res = resolve(resources, self._resources)
        h = C.c_void_p()
        _lib.check(_lib.lib().dyng_hip_result_clone(self._handle, res.stream_handle, C.byref(h)))
        return Result._wrap(h, res)

    @classmethod
    def from_arrays(cls, graph: Graph, source: int, distances, parents, *, canonicalize: bool = True,
                    options: Options | None = None, resources: Resources | None = None,
                    **kwargs: Any) -> "Result":
        """result::from_arrays (sssp.cpp:1118-1198): import, canonicalize and validate a tree."""
        _check_graph(graph)
        opt = dataclasses.replace(options or Options(), **kwargs)
        res = resolve(resources, graph.resources)
        n = graph.num_vertices
        d = np.asarray(distances, dtype=np.int64).reshape(-1)
        p = np.asarray(parents, dtype=np.int64).reshape(-1)
        if len(d) != n:
            raise InvalidArgumentError(f"sssp::result::from_arrays: {len(d)} distances for {n} vertices")
        if len(p) != n:
            raise InvalidArgumentError(f"sssp::result::from_arrays: {len(p)} parents for {n} vertices")
        if len(p) and (p.min() < -1 or p.max() >= n):
            bad = int(np.nonzero((p < -1) | (p >= n))[0][0])
            raise InvalidArgumentError(f"sssp::result::from_arrays: vertex {bad} has parent {int(p[bad])}, "
                                       f"outside [-1, {n})")
        d = np.ascontiguousarray(d)
        p32 = np.ascontiguousarray(p.astype(np.int32))
        h = C.c_void_p()
        o = opt._native()
        _lib.check(_lib.lib().dyng_hip_api_from_arrays(graph._handle, int(source), d.ctypes.data,
                                                       p32.ctypes.data, int(bool(canonicalize)), C.byref(o),
                                                       res.stream_handle, C.byref(h)))
        return cls._wrap(h, res)

    def differences(self, other: "Result") -> int:
        """Vertices (counted per array) where this tree and ``other`` differ, compared on the device."""
        if not isinstance(other, Result):
            raise TypeError("dyng_amd.sssp.Result.differences: other must be a Result")
        out = C.c_int64()
        _lib.check(_lib.lib().dyng_hip_result_equal(self._handle, other._handle, self._resources.stream_handle,
                                                    C.byref(out)))
        return int(out.value)

    def __reduce__(self):
        raise TypeError("dyng_amd.sssp.Result cannot be pickled; rebuild it with Result.from_arrays()")

    def __repr__(self) -> str:
        _, _, n = self.device_arrays()
        return f"dyng_amd.sssp.Result(source={self.source}, num_vertices={n}, graph_version={self.graph_version})"


def _check_graph(graph: Graph) -> None:
    if not isinstance(graph, Graph):
        raise TypeError("expected a dyng_amd.Graph")


def compute(graph: Graph, source: int, *, options: Options | None = None, resources: Resources | None = None,
            tuning: Tuning | None = None, **kwargs: Any) -> Result:
    """The canonical shortest-path tree of ``graph`` from ``source`` (sssp::compute,
    sssp.hpp:415-421): the Dijkstra tree with lowest-id ties."""
    _check_graph(graph)
    for key in kwargs:
        if key not in {f.name for f in dataclasses.fields(Options)}:
            raise TypeError(f"sssp.compute: unexpected keyword argument {key!r}")
    opt = dataclasses.replace(options or Options(), **kwargs)
    res = resolve(resources, graph.resources)
    o = opt._native()
    h = C.c_void_p()
    _lib.check(_lib.lib().dyng_hip_api_compute(graph._handle, int(source), C.byref(o), res.stream_handle,
                                               C.byref(h)))
    r = Result._wrap(h, res)
    if tuning is not None:
        r.set_tuning(tuning)
    return r


def compute_many(graph: Graph, sources, *, options: Options | None = None, resources: Resources | None = None,
                 tuning: Tuning | None = None, **kwargs: Any) -> Result:
    """The canonical trees of several distinct sources in one result and one launch (no
    counterpart in the reference: the batched path, e.g. one source per component of a disjoint
    union of graphs). Vertices reachable from several sources take the nearest one, lowest
    (distance, parent) first; for a disjoint union each component holds its own tree."""
    _check_graph(graph)
    for key in kwargs:
        if key not in {f.name for f in dataclasses.fields(Options)}:
            raise TypeError(f"sssp.compute_many: unexpected keyword argument {key!r}")
    opt = dataclasses.replace(options or Options(), **kwargs)
    res = resolve(resources, graph.resources)
    src = np.ascontiguousarray(np.asarray(sources, dtype=np.int64).reshape(-1))
    if src.size and (src.min() < -(2**31) or src.max() >= 2**31):
        raise InvalidArgumentError("sssp.compute_many: sources must fit int32")
    src32 = np.ascontiguousarray(src.astype(np.int32))
    o = opt._native()
    h = C.c_void_p()
    _lib.check(_lib.lib().dyng_hip_api_compute_multi(graph._handle, src32.ctypes.data, len(src32), C.byref(o),
                                                     res.stream_handle, C.byref(h)))
    r = Result._wrap(h, res)
    if tuning is not None:
        r.set_tuning(tuning)
    return r


def repair(graph: Graph, result: Result, *, resources: Resources | None = None) -> Stats:
    """The repair half of update (sssp_problem::resume + the engine, sssp.cpp:691-823) for a
    result exactly one ``Graph.apply`` behind the graph: ``graph.apply(batch)`` then
    ``repair(graph, result)`` equals ``update(graph, batch, result)``; with clones of the result,
    several schedules can be timed on one batch."""
    _check_graph(graph)
    if not isinstance(result, Result):
        raise TypeError("dyng_amd.sssp.repair: result must be a dyng_amd.sssp.Result")
    res = resolve(resources, result._resources)
    out = _lib.Stats()
    _lib.check(_lib.lib().dyng_hip_api_repair(graph._handle, result._handle, res.stream_handle, C.byref(out)))
    return Stats._from_native(out)


def update(graph: Graph, batch: EdgeBatch, result: Result, *, resources: Resources | None = None) -> Stats:
    """Apply ``batch`` to ``graph`` and update ``result`` (sssp::update, sssp.hpp:484-489)."""
    _check_graph(graph)
    if not isinstance(result, Result):
        raise TypeError("dyng_amd.sssp.update: result must be a dyng_amd.sssp.Result")
    if not isinstance(batch, EdgeBatch):
        raise TypeError("dyng_amd.sssp.update: batch must be a dyng_amd.EdgeBatch")
    res = resolve(resources, graph.resources)
    nb, keep = batch._native(graph.num_weights)
    st = _lib.Stats()
    _lib.check(_lib.lib().dyng_hip_api_update(graph._handle, C.byref(nb), result._handle, res.stream_handle,
                                              C.byref(st)))
    del keep
    return Stats._from_native(st)


# a stale result is reported by the library; the name is exported for callers
StaleResultError = StaleResultError
