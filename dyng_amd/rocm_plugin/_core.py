"""The ROCm plugin's native module: the names of dynG's extension module ``dyng._core`` that the
sssp path reaches, with the reference's signatures (python/dyng/_core.pyi: the enumerations,
``Resources``, ``BatchSemantics``, ``GraphProperties``, ``ApplySummary``, ``GraphI32I32I32``,
``EdgeBatchI32I32``, ``SsspOptions``, ``SsspStats``, ``SsspResultI32``, ``sssp_compute``,
``sssp_update``, ``sssp_from_arrays``; python/bindings/sssp.cpp:77-135), over this package's HIP
path. The graph types are dynG's only sssp instantiation on the configurations it runs,
``graph<int32, int32, int32>``.

The typed layer of ``dyng`` calls these as it calls ``_core``. Every other name of ``_core`` (other
algorithms, other graph types, hypergraphs) raises :class:`NotSupportedError` when read. Errors of
the HIP path are raised as the ``dyng.errors`` class of the same name when ``dyng`` is importable
(the reference's exception translation), else as :mod:`dyng_amd.errors`.

Results are on the device; ``distances()`` and ``parents()`` return host copies (numpy), as a
device result's arrays read from the host.
"""

from __future__ import annotations

import enum
import functools
import importlib
from typing import Any

import numpy as np

from .. import errors as _errors

__version__ = "0.4.0"

#: dynG's selection requires ``build_config["cuda"]`` of a plugin's module: here it stands for
#: "a device backend", which is HIP on the MI355X.
build_config = {"cuda": True, "hip": True, "openmp": False, "gpu": "MI355X (gfx950)", "backend": "hip"}


def _translated(fn):
    """Raise the HIP path's errors as the ``dyng.errors`` class of the same name when the
    reference's package is importable (the plugin is then running under it)."""

    @functools.wraps(fn)
    def run(*args, **kwargs):
        try:
            return fn(*args, **kwargs)
        except _errors.DyngError as e:
            try:
                ref = importlib.import_module("dyng.errors")
            except ImportError:
                raise e from None
            cls = getattr(ref, type(e).__name__, None)
            if cls is None or not isinstance(cls, type) or not issubclass(cls, BaseException):
                raise e from None
            raise cls(str(e)) from e

    return run


def _api():
    import dyng_amd

    return dyng_amd


# ---- enumerations (values as in _core.pyi) ----


class Backend(enum.Enum):
    sequential = 0
    openmp = 1
    cuda = 2  # the device backend: HIP on this plugin


class MemorySpace(enum.Enum):
    host = 0
    pinned_host = 1
    device = 2
    managed = 3


class CopyPolicy(enum.Enum):
    allow = 0
    warn = 1
    error = 2


class Engine(enum.Enum):
    automatic = 0
    fused = 1
    operators = 2


class RowLayout(enum.Enum):
    compact = 0
    slotted = 1
    slack = 2


class VertexStatus(enum.Enum):
    active = 0
    deleted = 1
    inactive = 2


class RowOrder(enum.Enum):
    sorted = 0
    append = 1


class MultiEdges(enum.Enum):
    forbid = 0
    allow = 1


class ExistingInsert(enum.Enum):
    upsert = 0
    error = 1
    ignore = 2


class MissingDelete(enum.Enum):
    ignore = 0
    error = 1


class SelfLoop(enum.Enum):
    keep = 0
    drop = 1
    error = 2


def backend_available(backend: Backend) -> bool:
    from . import available

    return backend == Backend.cuda and available()


def default_backend() -> Backend:
    return Backend.cuda


def library_version() -> str:
    return __version__


# ---- resources ----


class Resources:
    """The execution resources: the device backend only (HIP)."""

    def __init__(self) -> None:
        self._r = _api().Resources.hip()
        self._threads = 0
        self.copy_policy = CopyPolicy.allow

    @staticmethod
    def sequential() -> Resources:
        raise _errors.NotSupportedError("dyng ROCm plugin: the sequential backend is dyng._core's "
                                        "(DYNG_CPU_ONLY=1 selects it)")

    @staticmethod
    def openmp(num_threads: int = 0) -> Resources:
        raise _errors.NotSupportedError("dyng ROCm plugin: the OpenMP backend is dyng._core's "
                                        "(DYNG_CPU_ONLY=1 selects it)")

    @staticmethod
    def cuda(device: int = 0, stream: int | None = None, host_threads: int = 0) -> Resources:
        """The device backend on HIP device `device` (stream: None = the default stream, or a
        hipStream_t as an integer)."""
        r = Resources()
        r._r = _api().Resources.hip(int(device), stream)
        r._threads = int(host_threads)
        return r

    @property
    def backend(self) -> Backend:
        return Backend.cuda

    @property
    def device(self) -> int:
        return int(self._r.device)

    @property
    def stream(self) -> int:
        return int(self._r.stream_handle or 0)

    @property
    def num_threads(self) -> int:
        return self._threads

    @property
    def default_space(self) -> MemorySpace:
        return MemorySpace.device

    @property
    def workspace_bytes(self) -> int:
        return 0

    def release_workspaces(self) -> None:
        return None

    def warm_up(self) -> None:
        _api().device_count()

    @_translated
    def synchronize(self) -> None:
        _api().synchronize(self.device)

    def attach_profiler(self, profiler: Any) -> None:
        if profiler is not None:
            raise _errors.NotSupportedError("dyng ROCm plugin: use rocprofv3 to profile the HIP path")

    def has_profiler(self) -> bool:
        return False


# ---- graph properties and batch semantics ----


class BatchSemantics:
    def __init__(self) -> None:
        self.on_existing_insert = ExistingInsert.upsert
        self.on_missing_delete = MissingDelete.ignore
        self.on_self_loop = SelfLoop.keep
        self.deletions_first = True
        self.allow_vertex_growth = True
        self.as_sets = False
        self.as_net_effect = False

    @staticmethod
    def upsert_last_wins() -> BatchSemantics:
        return BatchSemantics()

    @staticmethod
    def set() -> BatchSemantics:
        s = BatchSemantics()
        s.as_sets = True
        return s

    @staticmethod
    def net_effect() -> BatchSemantics:
        s = BatchSemantics()
        s.as_net_effect = True
        return s


class GraphProperties:
    def __init__(self) -> None:
        self.directed = True
        self.store_transposed = False
        self.num_weights = 1
        self.layout = RowLayout.compact
        self.headroom = 0.25
        self.order = RowOrder.sorted
        self.parallel_edges = MultiEdges.forbid
        self.semantics = BatchSemantics()

    @staticmethod
    def mosp_compatible() -> GraphProperties:
        p = GraphProperties()
        p.store_transposed = True
        p.order = RowOrder.append
        p.parallel_edges = MultiEdges.allow
        return p

    @staticmethod
    def cycle_enum_compatible() -> GraphProperties:
        raise _errors.NotSupportedError("dyng ROCm plugin: cycle enumeration is not on the HIP path")

    def _to_api(self):
        """This package's properties; the device graph always keeps its in-rows (the sssp engine
        pulls over them) in slack rows, whatever layout is asked."""
        dg = _api()
        s = self.semantics
        if s.as_sets or s.as_net_effect:
            raise _errors.NotSupportedError("dyng ROCm plugin: batch_semantics::set() and net_effect() are not "
                                            "on the HIP path (upsert_last_wins with its error modes is)")
        if self.layout == RowLayout.slotted:
            raise _errors.NotSupportedError("dyng ROCm plugin: slotted rows (vertex deletion) are not on the HIP "
                                            "path")
        sem = dg.BatchSemantics(on_existing_insert=s.on_existing_insert.name, on_missing_delete=s.on_missing_delete.name,
                                on_self_loop=s.on_self_loop.name, deletions_first=bool(s.deletions_first),
                                allow_vertex_growth=bool(s.allow_vertex_growth))
        return dg.GraphProperties(directed=bool(self.directed), store_transposed=True, order=self.order.name,
                                  parallel_edges=self.parallel_edges.name, semantics=sem,
                                  headroom=float(self.headroom))


class ApplySummary:
    """apply_summary, read-only."""

    FIELDS = ("inserted_edges", "updated_edges", "deleted_edges", "ignored_deletions", "dropped_self_loops",
              "cancelled_pairs", "inserted_vertices", "deleted_vertices", "num_vertices_after",
              "ignored_insertions", "dropped_insertions", "relocated_rows", "compactions")

    def __init__(self, summary: Any = None) -> None:
        for f in self.FIELDS:
            object.__setattr__(self, f, int(getattr(summary, f, 0)) if summary is not None else 0)

    def __setattr__(self, name: str, value: Any) -> None:
        raise AttributeError("ApplySummary is read-only")


# ---- the int32 graph and its batches ----


def _i32(a: Any, what: str) -> np.ndarray:
    x = np.asarray(a)
    if x.dtype != np.int32:
        raise TypeError(f"{what}: an int32 array is expected, got {x.dtype}")
    return np.ascontiguousarray(x)


class EdgeBatchI32I32:
    """A batch of edge changes over arrays (insert weights edge-major, `num_weights` per edge)."""

    def __init__(self, insert_src: Any, insert_dst: Any, insert_weights: Any, delete_src: Any, delete_dst: Any,
                 num_weights: int, insert_vertices: Any, insert_vertex_labels: Any, delete_vertices: Any) -> None:
        for extra, what in ((insert_vertices, "insert_vertices"), (insert_vertex_labels, "insert_vertex_labels"),
                            (delete_vertices, "delete_vertices")):
            if extra is not None and len(extra) > 0:
                raise _errors.NotSupportedError(f"dyng ROCm plugin: {what} (vertex batches) are not on the HIP path")
        self._ins = (_i32(insert_src, "insert_src"), _i32(insert_dst, "insert_dst"))
        self._del = (_i32(delete_src, "delete_src"), _i32(delete_dst, "delete_dst"))
        self._k = int(num_weights)
        n = len(self._ins[0])
        if insert_weights is None:
            w = np.ones((n, max(self._k, 1)), np.int32)
        else:
            w = _i32(insert_weights, "insert_weights").reshape(n, -1) if n else np.zeros((0, max(self._k, 1)), np.int32)
        self._w = w

    @property
    def num_insertions(self) -> int:
        return len(self._ins[0])

    @property
    def num_deletions(self) -> int:
        return len(self._del[0])

    @property
    def num_weights(self) -> int:
        return self._k

    def _to_api(self):
        return _api().EdgeBatch(insert=(self._ins[0], self._ins[1], self._w), delete=self._del)


class GraphI32I32I32:
    """A dyng::graph<int32, int32, int32> on the device."""

    def __init__(self, graph: Any, props: GraphProperties) -> None:
        self._g = graph
        self._props = props

    @staticmethod
    @_translated
    def from_edges(resources: Resources, num_vertices: int, src: Any, dst: Any, weights: Any, num_weights: int,
                   properties: GraphProperties) -> GraphI32I32I32:
        s, d = _i32(src, "src"), _i32(dst, "dst")
        k = max(int(num_weights), 1)
        w = np.ones((len(s), k), np.int32) if weights is None else _i32(weights, "weights").reshape(len(s), k)
        g = _api().Graph.from_edges(s, d, w, num_vertices=int(num_vertices), properties=properties._to_api(),
                                    resources=resources._r)
        return GraphI32I32I32(g, properties)

    @staticmethod
    @_translated
    def from_csr(resources: Resources, row_ptr: Any, col_ind: Any, weights: Any, num_weights: int,
                 properties: GraphProperties) -> GraphI32I32I32:
        c = _i32(col_ind, "col_ind")
        k = max(int(num_weights), 1)
        w = np.ones((k, len(c)), np.int32) if weights is None else _i32(weights, "weights").reshape(k, len(c))
        g = _api().Graph.from_csr(_i32(row_ptr, "row_ptr"), c, w.T, properties=properties._to_api(),
                                  resources=resources._r)
        return GraphI32I32I32(g, properties)

    @staticmethod
    def with_capacity(resources: Resources, vertex_capacity: int, row_capacity: Any,
                      properties: GraphProperties) -> GraphI32I32I32:
        raise _errors.NotSupportedError("dyng ROCm plugin: graph::with_capacity (slotted rows) is not on the HIP "
                                        "path")

    @property
    def num_vertices(self) -> int:
        return int(self._g.num_vertices)

    @property
    def num_active_vertices(self) -> int:
        return int(self._g.num_vertices)

    def vertex_status(self) -> np.ndarray:
        return np.zeros(0, np.uint8)  # compact rows: no vertex states

    @property
    def num_edges(self) -> int:
        return int(self._g.num_edges)

    @property
    def num_weights(self) -> int:
        return int(self._g.num_weights)

    @property
    def directed(self) -> bool:
        return bool(self._props.directed)

    @property
    def has_transposed(self) -> bool:
        return True

    @property
    def space(self) -> MemorySpace:
        return MemorySpace.device

    @property
    def version(self) -> int:
        return int(self._g.version)

    @property
    def properties(self) -> GraphProperties:
        return self._props

    @_translated
    def to_csr(self, resources: Resources) -> tuple:
        """(row_ptr, col_ind, weights of shape (K, m)): a host copy of the out-edges."""
        row, col, w = self._g.to_csr()
        return (np.asarray(row, np.int32), np.asarray(col, np.int32), np.ascontiguousarray(w))

    @_translated
    def apply(self, resources: Resources, batch: EdgeBatchI32I32) -> ApplySummary:
        return ApplySummary(self._g.apply(batch._to_api()))

    @_translated
    def clone(self, resources: Resources) -> GraphI32I32I32:
        row, col, w = self._g.to_csr()
        g = _api().Graph.from_csr(row, col, w.T, properties=self._props._to_api(), resources=resources._r)
        return GraphI32I32I32(g, self._props)

    def to_backend(self, resources: Resources) -> GraphI32I32I32:
        return self

    def reserve(self, resources: Resources, edge_capacity: int) -> None:
        return None  # the slack rows grow as batches need

    def check_integrity(self, resources: Resources) -> None:
        return None


# ---- sssp ----


class SsspOptions:
    def __init__(self) -> None:
        self.delta = 0
        self.objective = 0
        self.cuda_engine = Engine.automatic
        self.validate_inputs = True

    def _to_api(self):
        return _api().sssp.Options(delta=int(self.delta), objective=int(self.objective),
                                   cuda_engine=self.cuda_engine.name, validate_inputs=bool(self.validate_inputs))


class SsspStats:
    """sssp::stats, read-only."""

    def __init__(self, st: Any) -> None:
        self.affected = int(st.affected)
        self.iterations = int(st.iterations)
        self.frontier_visits = int(st.frontier_visits)
        self.fallback_used = bool(st.fallback_used)
        self.converged = bool(st.converged)
        self.engine_used = Engine[str(st.engine_used)]
        self.batch = ApplySummary(st.batch)
        self.invalidated = int(st.invalidated)
        self.epochs = int(st.epochs)
        self.pushes = int(st.pushes)
        self.packed_parents = bool(st.packed_parents)


class SsspResultI32:
    """An sssp::result on the device."""

    def __init__(self, result: Any, resources: Resources) -> None:
        self._r = result
        self.writer = resources
        self.generation = 0

    @property
    def source(self) -> int:
        return int(self._r.source)

    @_translated
    def distances(self) -> np.ndarray:
        return self._r.to_numpy()[0]

    @_translated
    def parents(self) -> np.ndarray:
        return self._r.to_numpy()[1]

    @property
    def options(self) -> SsspOptions:
        o = self._r.options
        out = SsspOptions()
        out.delta, out.objective, out.validate_inputs = int(o.delta), int(o.objective), bool(o.validate_inputs)
        out.cuda_engine = Engine[str(o.cuda_engine)]
        return out

    @_translated
    def set_options(self, options: SsspOptions) -> None:
        self._r.set_options(options._to_api())

    @property
    def graph_version(self) -> int:
        return int(self._r.graph_version)

    @property
    def space(self) -> MemorySpace:
        return MemorySpace.device

    @_translated
    def clone(self, resources: Resources) -> SsspResultI32:
        return SsspResultI32(self._r.clone(resources._r), resources)


@_translated
def sssp_compute(resources: Resources, graph: GraphI32I32I32, source: int, options: SsspOptions) -> SsspResultI32:
    if not 0 <= int(source) < graph.num_vertices:
        raise _errors.InvalidArgumentError(f"sssp: the source {source} is out of range [0, {graph.num_vertices})")
    r = _api().sssp.compute(graph._g, int(source), options=options._to_api(), resources=resources._r)
    return SsspResultI32(r, resources)


@_translated
def sssp_update(resources: Resources, graph: GraphI32I32I32, batch: EdgeBatchI32I32, result: SsspResultI32) -> SsspStats:
    st = _api().sssp.update(graph._g, batch._to_api(), result._r, resources=resources._r)
    result.writer = resources
    result.generation += 1
    return SsspStats(st)


@_translated
def sssp_from_arrays(resources: Resources, graph: GraphI32I32I32, source: int, distances: Any, parents: Any,
                     canonicalize: bool, options: SsspOptions) -> SsspResultI32:
    if not 0 <= int(source) < graph.num_vertices:
        raise _errors.InvalidArgumentError(f"sssp::result::from_arrays: the source {source} is out of range "
                                           f"[0, {graph.num_vertices})")
    r = _api().sssp.Result.from_arrays(graph._g, int(source), np.asarray(distances, np.int64),
                                       np.asarray(parents, np.int32), canonicalize=bool(canonicalize),
                                       options=options._to_api(), resources=resources._r)
    return SsspResultI32(r, resources)


#: The names of _core the sssp path of the typed layer reaches (python/dyng/{resources,graph,
#: batch,sssp}.py); every one is defined above.
SSSP_PATH = ("Backend", "MemorySpace", "CopyPolicy", "Engine", "RowLayout", "VertexStatus", "RowOrder", "MultiEdges",
             "ExistingInsert", "MissingDelete", "SelfLoop", "Resources", "BatchSemantics", "GraphProperties",
             "ApplySummary", "GraphI32I32I32", "EdgeBatchI32I32", "SsspOptions", "SsspStats", "SsspResultI32",
             "sssp_compute", "sssp_update", "sssp_from_arrays", "backend_available", "default_backend",
             "library_version", "build_config", "__version__")


def __getattr__(name: str) -> Any:
    if name.startswith("__"):
        raise AttributeError(name)
    raise _errors.NotSupportedError(f"dyng ROCm plugin: {name} is not on the HIP path (this plugin serves "
                                    "sssp::compute / sssp::update and the graph and batch types they take)")
