"""dyng_amd.Graph: a graph resident on the device (python/dyng/graph.py; C++ dyng::graph).

The device copy is the graph: slack rows kept sorted by neighbour id (parallel edges in insertion
order), out-rows and in-rows, updated in place by each batch. `to_csr()` reads the rows back.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .errors import InvalidArgumentError, NotSupportedError
from .resources import Resources, resolve


@dataclass
class BatchSemantics:
    """batch_semantics (cpp/include/dyng/graph/graph_properties.hpp:67-188)."""

    on_existing_insert: str = "upsert"  # upsert | error | ignore
    on_missing_delete: str = "ignore"  # ignore | error
    on_self_loop: str = "keep"  # keep | drop | error
    deletions_first: bool = True
    allow_vertex_growth: bool = True

    @staticmethod
    def upsert_last_wins() -> "BatchSemantics":
        return BatchSemantics()


@dataclass
class GraphProperties:
    """graph_properties (graph_properties.hpp:195-241)."""

    directed: bool = True
    store_transposed: bool = True
    order: str = "sorted"  # sorted | append
    parallel_edges: str = "forbid"  # forbid | allow
    semantics: BatchSemantics = field(default_factory=BatchSemantics)
    headroom: float = 0.25

    @staticmethod
    def mosp_compatible() -> "GraphProperties":
        return GraphProperties(order="append", parallel_edges="allow")

    def _native(self) -> _lib.GraphProps:
        s = self.semantics
        ins = {"upsert": 0, "error": 1, "ignore": 2}
        dele = {"ignore": 0, "error": 1}
        loops = {"keep": 0, "drop": 1, "error": 2}
        try:
            return _lib.GraphProps(int(self.directed), int(self.order == "sorted"),
                                   int(self.parallel_edges == "allow"), ins[s.on_existing_insert],
                                   dele[s.on_missing_delete], loops[s.on_self_loop],
                                   int(s.deletions_first), int(s.allow_vertex_growth),
                                   float(self.headroom))
        except KeyError as e:
            raise InvalidArgumentError(f"graph properties: unknown rule {e}") from None


@dataclass(frozen=True)
class ApplySummary:
    """apply_summary (cpp/include/dyng/graph/apply_summary.hpp:22-41)."""

    inserted_edges: int = 0
    updated_edges: int = 0
    deleted_edges: int = 0
    ignored_deletions: int = 0
    dropped_self_loops: int = 0
    inserted_vertices: int = 0
    num_vertices_after: int = 0
    ignored_insertions: int = 0
    relocated_rows: int = 0
    compactions: int = 0
    host_plans: int = field(default=0, compare=False)  # merge plans read back before the merges
    stage_ms: dict = field(default_factory=dict, compare=False)  # device time per apply stage

    @classmethod
    def _from_native(cls, s: _lib.ApplySummary) -> "ApplySummary":
        stages = {name: float(s.stage_ms[i]) for i, name in enumerate(_lib.APPLY_STAGES)}
        return cls(**{f: int(getattr(s, f)) for f in _lib.SUMMARY_FIELDS}, stage_ms=stages)


def _ids(x, what: str) -> np.ndarray:
    a = np.asarray(x)
    if a.size and not np.issubdtype(a.dtype, np.integer):
        raise InvalidArgumentError(f"{what}: ids must be integers")
    a = a.astype(np.int64).reshape(-1)
    if a.size and (a.min() < -(2**31) or a.max() >= 2**31 - 1):
        raise NotSupportedError(f"{what}: the HIP backend takes int32 vertex ids")
    return np.ascontiguousarray(a.astype(np.int32))


def _weights(w, count: int, what: str) -> np.ndarray:
    if w is None:
        raise NotSupportedError(
            "dyng.sssp needs integer weights; this graph is unweighted (build it with weights=)")
    a = np.asarray(w)
    if a.size and not np.issubdtype(a.dtype, np.integer):
        raise NotSupportedError(f"{what}: sssp needs integer (int32) weights")
    a = a.astype(np.int64)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    if a.shape[0] != count:
        raise InvalidArgumentError(f"{what}: {a.shape[0]} weight rows for {count} edges")
    if a.size and (a.min() < -(2**31) or a.max() > 2**31 - 1):
        raise InvalidArgumentError(f"{what}: weights must fit int32")
    return np.ascontiguousarray(a.astype(np.int32))


class Graph:
    """A directed (or symmetric) weighted graph on the device."""

    __slots__ = ("_handle", "_resources", "_props", "_K")

    def __init__(self):
        raise TypeError("use dyng_amd.Graph.from_edges() or Graph.from_csr()")

    @classmethod
    def _wrap(cls, handle: C.c_void_p, res: Resources, props: GraphProperties, K: int) -> "Graph":
        self = object.__new__(cls)
        self._handle = handle
        self._resources = res
        self._props = props
        self._K = K
        return self

    @classmethod
    def from_edges(cls, src, dst, weights=None, *, num_vertices: int | None = None,
                   properties: GraphProperties | None = None, resources: Resources | None = None,
                   directed: bool | None = None, store_transposed: bool | None = None) -> "Graph":
        """Build from an edge list (build_from_edges_host, apply_host.cpp:608-690): ids in
        [0, num_vertices); under multi_edges::forbid repeated (u,v) keep the last weights."""
        props = properties or GraphProperties()
        if directed is not None:
            props.directed = bool(directed)
        if store_transposed is not None:
            props.store_transposed = bool(store_transposed)
        if not props.store_transposed:
            raise InvalidArgumentError(
                "dyng_amd: sssp needs the in-edges; build the graph with store_transposed = True")
        res = resolve(resources)
        s = _ids(src, "Graph.from_edges: src")
        d = _ids(dst, "Graph.from_edges: dst")
        if len(s) != len(d):
            raise InvalidArgumentError("Graph.from_edges: src and dst differ in length")
        w = _weights(weights, len(s), "Graph.from_edges: weights")
        n = int(num_vertices) if num_vertices is not None else (int(max(s.max(), d.max())) + 1 if len(s) else 0)
        K = w.shape[1] if w.ndim == 2 and w.size else (1 if w.ndim == 1 else w.shape[1])
        if len(s) and (min(s.min(), d.min()) < 0 or max(s.max(), d.max()) >= n):
            raise InvalidArgumentError(f"Graph.from_edges: an edge is out of range [0, {n})")
        handle = C.c_void_p()
        np_props = props._native()
        _lib.check(_lib.lib().dyng_hip_graph_create_from_edges(
            res.device, C.byref(np_props), n, len(s), K, s.ctypes.data, d.ctypes.data, w.ctypes.data,
            0, res.stream_handle, C.byref(handle)))
        return cls._wrap(handle, res, props, K)

    @classmethod
    def from_csr(cls, row_ptr, col, weights, *, properties: GraphProperties | None = None,
                 resources: Resources | None = None) -> "Graph":
        """Build from compact rows; weights are (m,) or (m, K) (or objective-major (K, m) with
        `weights.T`)."""
        props = properties or GraphProperties()
        res = resolve(resources)
        row = np.ascontiguousarray(np.asarray(row_ptr, dtype=np.int64))
        c = _ids(col, "Graph.from_csr: col")
        w = _weights(weights, len(c), "Graph.from_csr: weights")
        n = len(row) - 1
        K = w.shape[1]
        wt = np.ascontiguousarray(w.T)
        handle = C.c_void_p()
        np_props = props._native()
        _lib.check(_lib.lib().dyng_hip_graph_create(
            res.device, C.byref(np_props), n, len(c), K, row.ctypes.data, c.ctypes.data,
            wt.ctypes.data, 0, res.stream_handle, C.byref(handle)))
        return cls._wrap(handle, res, props, K)

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h and _lib._lib is not None:
            _lib._lib.dyng_hip_graph_destroy(h)
            self._handle = None

    @property
    def resources(self) -> Resources:
        return self._resources

    @property
    def properties(self) -> GraphProperties:
        return self._props

    @property
    def num_vertices(self) -> int:
        n = C.c_int64()
        _lib.check(_lib.lib().dyng_hip_graph_num_vertices(self._handle, C.byref(n)))
        return n.value

    @property
    def num_edges(self) -> int:
        m = C.c_int64()
        _lib.check(_lib.lib().dyng_hip_graph_num_edges(self._handle, C.byref(m)))
        return m.value

    @property
    def num_weights(self) -> int:
        return self._K

    @property
    def version(self) -> int:
        v = C.c_uint64()
        _lib.check(_lib.lib().dyng_hip_graph_version(self._handle, C.byref(v)))
        return v.value

    @property
    def vertex_dtype(self) -> np.dtype:
        return np.dtype(np.int32)

    @property
    def weight_dtype(self) -> np.dtype:
        return np.dtype(np.int32)

    def column(self, k: int = 0) -> tuple[int, int]:
        """(largest weight, weight sum) of objective k."""
        mx, sm = C.c_int64(), C.c_int64()
        _lib.check(_lib.lib().dyng_hip_graph_column(self._handle, k, C.byref(mx), C.byref(sm)))
        return mx.value, sm.value

    def _download(self, fn) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        n, m, K = self.num_vertices, self.num_edges, self._K
        row = np.zeros(n + 1, np.int64)
        col = np.zeros(max(m, 1), np.int32)
        w = np.zeros(max(m * K, 1), np.int32)
        _lib.check(fn(self._handle, row.ctypes.data, col.ctypes.data, w.ctypes.data))
        return row, col[:m], w[: m * K].reshape(K, m)

    def to_csr(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Compact out-rows (row_ptr, col, weights (K, m)), rows sorted by neighbour id."""
        return self._download(_lib.lib().dyng_hip_graph_download)

    def in_csr(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Compact in-rows (row_ptr, source ids, weights (K, m)), rows sorted by source id."""
        return self._download(_lib.lib().dyng_hip_graph_download_in)

    def apply(self, batch) -> ApplySummary:
        """Apply a batch to the graph alone (graph::apply); results of this graph go stale."""
        nb, keep = batch._native(self._K)
        out = _lib.ApplySummary()
        _lib.check(_lib.lib().dyng_hip_graph_apply(self._handle, C.byref(nb), self._resources.stream_handle,
                                                   C.byref(out)))
        del keep
        return ApplySummary._from_native(out)

    def __repr__(self) -> str:
        return (f"dyng_amd.Graph(num_vertices={self.num_vertices}, num_edges={self.num_edges}, "
                f"version={self.version})")
