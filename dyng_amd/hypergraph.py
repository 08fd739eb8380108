"""Weighted hypergraphs with slack_csr storage: the container dyng_amd.hyper_sssp repairs its
shortest hyperpaths on (python/dyng/hypergraph.py, cpp/include/dyng/hypergraph/hypergraph.hpp).

The container lives in libdyng_hip.so (``dyng_hip_hypergraph_*``, host memory, C++): this
module mirrors the reference's Python names over it.

- ``Hypergraph.from_lists(lists, weights, ...)``: hyperedge i gets id i, its members sorted and
  deduplicated; ``h2v()``, ``v2h()``, ``h2h()`` (the line graph: ordered pairs of live
  hyperedges that share a vertex), ``weights``, ``version``, ``hyperedge_id_bound``;
- ``apply(HyperedgeBatch)``: the phases of the reference's apply (hypergraph.cpp:613-760,
  hyperedge_batch.hpp:27-40): hyperedge deletions (a negative, out-of-range, dead or repeated id
  is skipped), incidence removals, incidence additions, hyperedge insertions (repeated members
  dropped). An edit is checked against the hypergraph before the batch, hyperedge by hyperedge in
  batch order, and skipped when its hyperedge is not live then or deleted by the batch, its
  vertex is out of range, a removal names a non-member or the last member, or an addition names a
  member. Every skip is counted in ``HyperApplySummary.skipped_ops``;
- the ids the insertions get: the store's best-fit id allocator (id_allocator.cpp:119-580) under
  the batch order of ``id_reuse`` (``erase_first``: deletions, removals, additions, insertions;
  ``insert_first``: insertions first) and the payload sizing, compaction and growth of the
  ``CapacityPolicy`` (capacity.cpp:105-170, store_policy.cpp:71-116).

The escher storage, unweighted hypergraphs' algorithms and the hypergraph file formats are not
built.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Any, Iterable, NamedTuple

import numpy as np

from . import _lib
from .errors import CapacityError, InvalidArgumentError, NotSupportedError

__all__ = ["CapacityError", "CapacityPolicy", "HypergraphProperties", "HyperApplySummary", "HyperedgeBatch",
           "Hypergraph", "Incidence", "LineGraph"]

INT_MAX = 2**31 - 1
MAX_WEIGHT = 2**28


@dataclass(frozen=True)
class CapacityPolicy:
    """How the storage is sized and grows (hypergraph_properties.hpp:82-140)."""

    initial_headroom: float = 0.125
    initial_slack: int = 1024
    min_initial_payload: int = 0
    grow_factor: float = 1.5
    compact_before_grow: bool = True
    max_payload: int = 0
    row_alignment: int = 4
    compaction_headroom: int = 8

    @classmethod
    def escher_gpu(cls, min_initial_payload: int = 0) -> "CapacityPolicy":
        return cls(0.125, 1024, int(min_initial_payload), 1.5, False, 0, 4, 0)

    @classmethod
    def mosp_escher(cls, extra: int = 1 << 20) -> "CapacityPolicy":
        return cls(0.3, int(extra), 0, 0.0, True, 2000000000, 4, 0)

    @classmethod
    def preset(cls, name: str) -> "CapacityPolicy":
        presets = {"dyng_default": cls(), "escher_gpu": cls.escher_gpu(), "mosp_escher": cls.mosp_escher()}
        if name not in presets:
            raise InvalidArgumentError(f"CapacityPolicy.preset: unknown preset {name!r}")
        return presets[name]

    def limit(self) -> int:
        """The largest payload (capacity.cpp:107-109)."""
        return min(self.max_payload, INT_MAX) if self.max_payload > 0 else INT_MAX

    def initial_payload(self, size: int) -> int:
        """The initial capacity of a tree of `size` entries (capacity.cpp:147-170)."""
        limit = self.limit()
        scaled = size * self.initial_headroom
        slack = int(scaled) if scaled < limit else limit
        cap = size + slack + min(self.initial_slack, limit)
        cap = min(max(cap, self.min_initial_payload), limit)
        if cap < size:
            raise CapacityError(f"dyng: hypergraph: a tree of {size} payload entries exceeds the largest payload "
                                f"of its capacity policy ({limit} entries)")
        return cap

    def _native(self, reuse: str) -> _lib.HyperProps:
        return _lib.HyperProps(1 if reuse == "insert_first" else 0, float(self.initial_headroom),
                               int(self.initial_slack), int(self.min_initial_payload), float(self.grow_factor),
                               int(bool(self.compact_before_grow)), int(self.max_payload), int(self.row_alignment),
                               int(self.compaction_headroom))


@dataclass(frozen=True)
class HypergraphProperties:
    """hypergraph_properties (hypergraph_properties.hpp:143-152): the storage, the line graph's
    upkeep, the order of a batch (which decides the reused ids), the sizing."""

    storage: str = "slack_csr"
    line_graph: str = "automatic"
    reuse: str = "erase_first"
    capacity: CapacityPolicy = field(default_factory=CapacityPolicy)
    store_v2h: bool = True


@dataclass(frozen=True)
class HyperApplySummary:
    """hyper_apply_summary (hyper_apply_summary.hpp:20-36)."""

    inserted_hyperedges: int = 0
    deleted_hyperedges: int = 0
    inserted_incidences: int = 0
    deleted_incidences: int = 0
    skipped_ops: int = 0
    ignored_deletions: int = 0
    dropped_members: int = 0
    num_hyperedges_after: int = 0
    id_bound_after: int = 0
    inserted_ids: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))


class Incidence(NamedTuple):
    """A CSR of rows (hyperedge -> members, or vertex -> hyperedges)."""

    row_ptr: np.ndarray
    col_ind: np.ndarray
    version: int

    def row(self, r: int) -> np.ndarray:
        return self.col_ind[self.row_ptr[r]:self.row_ptr[r + 1]]

    @property
    def num_rows(self) -> int:
        return len(self.row_ptr) - 1


class LineGraph(NamedTuple):
    """The line graph: per id, the live ids sharing a vertex with it, sorted."""

    row_ptr: np.ndarray
    col_ind: np.ndarray
    version: int

    def row(self, r: int) -> np.ndarray:
        return self.col_ind[self.row_ptr[r]:self.row_ptr[r + 1]]

    @property
    def num_rows(self) -> int:
        return len(self.row_ptr) - 1


def _csr(rows: list[Iterable[int]], version: int, cls: Any) -> Any:
    counts = [len(r) for r in rows]
    ptr = np.zeros(len(rows) + 1, np.int64)
    ptr[1:] = np.cumsum(counts)
    col = np.fromiter((v for r in rows for v in r), dtype=np.int64, count=int(ptr[-1]))
    return cls(ptr, col, version)


class HyperedgeBatch:
    """hyperedge_batch (hyperedge_batch.hpp:22-367): hyperedge deletions (ids), insertions
    (member lists, one weight each on a weighted hypergraph) and incidence edits ((hyperedge,
    vertex) removals and additions). The phases run in a fixed order whatever the order of the
    arrays (see :meth:`Hypergraph.apply`)."""

    def __init__(self, delete_hyperedges: Any = None, insert_hyperedges: Any = None, insert_weights: Any = None,
                 *, delete_incidences: Any = None, insert_incidences: Any = None,
                 insert_offsets: Any = None, insert_members: Any = None):
        self._delete = self._ids(delete_hyperedges, "delete_hyperedges")
        if insert_offsets is not None or insert_members is not None:
            if insert_hyperedges is not None:
                raise InvalidArgumentError("HyperedgeBatch: give insert_hyperedges or insert_offsets/insert_members")
            off = np.asarray(insert_offsets if insert_offsets is not None else [0], np.int64)
            mem = np.asarray(insert_members if insert_members is not None else [], np.int64)
            if len(off) == 0 or off[0] != 0 or off[-1] != len(mem) or np.any(np.diff(off) < 0):
                raise InvalidArgumentError("HyperedgeBatch: insert_offsets must start at 0, grow and end at the "
                                           "member count")
            self._insertions = [mem[off[i]:off[i + 1]].tolist() for i in range(len(off) - 1)]
        else:
            self._insertions = [list(np.asarray(h, np.int64).tolist()) for h in (insert_hyperedges or [])]
        if insert_weights is None:
            self._weights = None
        else:
            self._weights = np.asarray(insert_weights, np.int64).reshape(-1)
            if len(self._weights) != len(self._insertions):
                raise InvalidArgumentError(f"HyperedgeBatch: {len(self._weights)} insert_weights for "
                                           f"{len(self._insertions)} insertions")
        self._del_inc = self._pair(delete_incidences, "delete_incidences")
        self._ins_inc = self._pair(insert_incidences, "insert_incidences")

    @staticmethod
    def _ids(x: Any, what: str) -> np.ndarray:
        a = np.asarray([] if x is None else x)
        if a.size and not np.issubdtype(a.dtype, np.integer):
            raise InvalidArgumentError(f"HyperedgeBatch.{what}: expected integers")
        return a.astype(np.int64).reshape(-1)

    @classmethod
    def _pair(cls, value: Any, what: str) -> tuple[np.ndarray, np.ndarray]:
        if value is None:
            return np.zeros(0, np.int64), np.zeros(0, np.int64)
        h, v = value
        h, v = cls._ids(h, what), cls._ids(v, what)
        if len(h) != len(v):
            raise InvalidArgumentError(f"HyperedgeBatch.{what}: {len(h)} hyperedges for {len(v)} vertices")
        return h, v

    @property
    def delete_hyperedges(self) -> np.ndarray:
        return self._delete.copy()

    @property
    def insert_offsets(self) -> np.ndarray:
        off = np.zeros(len(self._insertions) + 1, np.int64)
        off[1:] = np.cumsum([len(h) for h in self._insertions])
        return off

    @property
    def insert_members(self) -> np.ndarray:
        return np.asarray([v for h in self._insertions for v in h], np.int64)

    @property
    def insert_weights(self) -> np.ndarray | None:
        return None if self._weights is None else self._weights.copy()

    @property
    def delete_incidences(self) -> tuple[np.ndarray, np.ndarray]:
        return self._del_inc[0].copy(), self._del_inc[1].copy()

    @property
    def insert_incidences(self) -> tuple[np.ndarray, np.ndarray]:
        return self._ins_inc[0].copy(), self._ins_inc[1].copy()

    @property
    def num_incidence_edits(self) -> int:
        return len(self._del_inc[0]) + len(self._ins_inc[0])

    @property
    def num_insertions(self) -> int:
        return len(self._insertions)

    @property
    def num_deletions(self) -> int:
        return len(self._delete)

    @property
    def empty(self) -> bool:
        return self.num_insertions == 0 and self.num_deletions == 0 and self.num_incidence_edits == 0

    def insertions(self) -> list[list[int]]:
        return [list(h) for h in self._insertions]

    def __repr__(self) -> str:
        return (f"dyng_amd.HyperedgeBatch(deletions={self.num_deletions}, insertions={self.num_insertions}, "
                f"incidence_edits={self.num_incidence_edits})")




def _ptr(a: np.ndarray) -> Any:
    return a.ctypes.data if a.size else None


class Hypergraph:
    """A weighted hypergraph with slack_csr storage (hypergraph.hpp; python/dyng/hypergraph.py),
    over the library's host container (``dyng_hip_hypergraph``)."""

    __slots__ = ("_h", "_props", "__weakref__")

    def __init__(self) -> None:
        raise TypeError("use dyng_amd.Hypergraph.from_lists()")

    @classmethod
    def from_lists(cls, lists: Iterable[Iterable[int]], weights: Iterable[int] | None = None, *,
                   num_vertices: int | None = None, storage: str = "slack_csr", line_graph: str = "automatic",
                   id_reuse: str = "erase_first", capacity: CapacityPolicy | str = "dyng_default",
                   store_v2h: bool = True, resources: Any = None) -> "Hypergraph":
        """Hyperedge i gets id i; repeated members are dropped (hypergraph::from_lists)."""
        if storage != "slack_csr":
            raise NotSupportedError(f"dyng_amd.Hypergraph: storage {storage!r} is not built (slack_csr only)")
        if line_graph not in ("automatic", "slack_csr"):
            raise NotSupportedError(f"dyng_amd.Hypergraph: line_graph {line_graph!r} is not built with slack_csr "
                                    "storage (automatic or slack_csr)")
        if id_reuse not in ("erase_first", "insert_first"):
            raise InvalidArgumentError(f"dyng_amd.Hypergraph: id_reuse must be 'erase_first' or 'insert_first', "
                                       f"got {id_reuse!r}")
        if not store_v2h:
            raise InvalidArgumentError("dyng_amd.Hypergraph: slack_csr storage keeps store_v2h")
        policy = CapacityPolicy.preset(capacity) if isinstance(capacity, str) else capacity
        raw = [np.asarray(list(h), np.int64).reshape(-1) for h in lists]
        off = np.zeros(len(raw) + 1, np.int64)
        off[1:] = np.cumsum([len(h) for h in raw])
        mem64 = np.concatenate(raw) if raw else np.zeros(0, np.int64)
        n = int(num_vertices) if num_vertices is not None else (int(mem64.max()) + 1 if mem64.size else 0)
        if mem64.size and (mem64.min() < 0 or mem64.max() >= n):
            raise InvalidArgumentError(f"dyng_amd.Hypergraph.from_lists: a member is outside [0, {n})")
        mem = np.ascontiguousarray(mem64.astype(np.int32))
        w = None
        if weights is not None:
            w = np.ascontiguousarray(np.asarray(list(weights), np.int64).reshape(-1))
            if len(w) != len(raw):
                raise InvalidArgumentError(f"dyng_amd.Hypergraph.from_lists: {len(w)} weights for {len(raw)} hyperedges")
        props = policy._native(id_reuse)
        h = C.c_void_p()
        _lib.check(_lib.lib().dyng_hip_hypergraph_create(off.ctypes.data, _ptr(mem), len(raw),
                                                         None if w is None else (w.ctypes.data if w.size else
                                                                                 np.zeros(1, np.int64).ctypes.data),
                                                         n, C.byref(props), C.byref(h)))
        self = object.__new__(cls)
        self._h = h
        self._props = HypergraphProperties("slack_csr", line_graph, id_reuse, policy, True)
        return self

    def __del__(self) -> None:
        h = getattr(self, "_h", None)
        if h and _lib._lib is not None:
            _lib._lib.dyng_hip_hypergraph_destroy(h)
            self._h = None

    def _info(self) -> tuple[int, int, int, int, bool]:
        n, live, bound, version, weighted = C.c_int32(), C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
        _lib.check(_lib.lib().dyng_hip_hypergraph_info(self._h, C.byref(n), C.byref(live), C.byref(bound),
                                                       C.byref(version), C.byref(weighted)))
        return n.value, live.value, bound.value, version.value, bool(weighted.value)

    # ---- views ----

    @property
    def num_vertices(self) -> int:
        return self._info()[0]

    @property
    def num_hyperedges(self) -> int:
        return self._info()[1]

    @property
    def hyperedge_id_bound(self) -> int:
        return self._info()[2]

    @property
    def version(self) -> int:
        return self._info()[3]

    @property
    def properties(self) -> HypergraphProperties:
        return self._props

    @property
    def weighted(self) -> bool:
        return self._info()[4]

    @property
    def space(self) -> str:
        return "host"

    def _rows_csr(self) -> tuple[np.ndarray, np.ndarray, np.ndarray | None]:
        off, mem, w = _lib._i64p(), _lib._i32p(), _lib._i64p()
        _lib.check(_lib.lib().dyng_hip_hypergraph_rows(self._h, C.byref(off), C.byref(mem), C.byref(w)))
        b = self.hyperedge_id_bound
        offsets = np.ctypeslib.as_array(off, shape=(b + 1,)).copy()
        total = int(offsets[-1])
        members = np.ctypeslib.as_array(mem, shape=(total,)).copy() if total else np.zeros(0, np.int32)
        weights = np.ctypeslib.as_array(w, shape=(b,)).copy() if (w and b) else (np.zeros(0, np.int64) if w else None)
        return offsets, members, weights

    @property
    def weights(self) -> np.ndarray | None:
        """The weight of every id below the bound (0 for a dead id)."""
        if not self.weighted:
            return None
        return self._rows_csr()[2]

    def rows(self) -> list[list[int]]:
        """The members of every id below the bound (empty for a dead id)."""
        off, mem, _ = self._rows_csr()
        return [mem[off[i]:off[i + 1]].tolist() for i in range(len(off) - 1)]

    def h2v(self) -> Incidence:
        off, mem, _ = self._rows_csr()
        return Incidence(off, mem.astype(np.int64), self.version)

    def v2h(self) -> Incidence:
        rows = self.rows()
        by_v: list[list[int]] = [[] for _ in range(self.num_vertices)]
        for h, r in enumerate(rows):
            for v in r:
                by_v[v].append(h)
        return _csr(by_v, self.version, Incidence)

    def h2h(self) -> LineGraph:
        """The line graph (ordered pairs of live ids that share a vertex, rows sorted)."""
        rows = self.rows()
        by_v: dict[int, list[int]] = {}
        for h, r in enumerate(rows):
            for v in r:
                by_v.setdefault(v, []).append(h)
        nb = [set() for _ in rows]
        for hs in by_v.values():
            for a in hs:
                nb[a].update(hs)
        return _csr([sorted(s - {h}) for h, s in enumerate(nb)], self.version, LineGraph)

    def clone(self) -> "Hypergraph":
        h = C.c_void_p()
        _lib.check(_lib.lib().dyng_hip_hypergraph_clone(self._h, C.byref(h)))
        out = object.__new__(Hypergraph)
        out._h = h
        out._props = self._props
        return out

    def __copy__(self) -> "Hypergraph":
        return self.clone()

    def __deepcopy__(self, memo: dict) -> "Hypergraph":
        return self.clone()

    def __repr__(self) -> str:
        n, live, bound, version, _ = self._info()
        return (f"dyng_amd.Hypergraph(num_vertices={n}, num_hyperedges={live}, hyperedge_id_bound={bound}, "
                f"version={version})")

    # ---- apply ----

    def apply(self, batch: HyperedgeBatch) -> HyperApplySummary:
        """Apply a batch (hypergraph::apply); the version grows by one."""
        return self._apply(batch)

    def _apply(self, batch: HyperedgeBatch, max_weight: int = 0) -> HyperApplySummary:
        if not isinstance(batch, HyperedgeBatch):
            raise TypeError("dyng_amd.Hypergraph.apply: expected a dyng_amd.HyperedgeBatch")
        dels = np.ascontiguousarray(batch._delete, np.int64)
        off = np.ascontiguousarray(batch.insert_offsets, np.int64)
        mem64 = batch.insert_members
        n = self.num_vertices
        if mem64.size and (mem64.min() < 0 or mem64.max() >= n):
            raise InvalidArgumentError(f"dyng_amd.Hypergraph.apply: a member of an inserted hyperedge is outside "
                                       f"[0, {n})")
        mem = np.ascontiguousarray(mem64.astype(np.int32))
        w = None if batch._weights is None else np.ascontiguousarray(batch._weights, np.int64)
        rh, rv = (np.ascontiguousarray(batch._del_inc[0], np.int64),
                  np.ascontiguousarray(np.clip(batch._del_inc[1], -1, INT_MAX).astype(np.int32)))
        ah, av = (np.ascontiguousarray(batch._ins_inc[0], np.int64),
                  np.ascontiguousarray(np.clip(batch._ins_inc[1], -1, INT_MAX).astype(np.int32)))
        nb = _lib.HyperBatch(_ptr(dels), len(dels), off.ctypes.data, _ptr(mem), None if w is None else _ptr(w),
                             batch.num_insertions, _ptr(rh), _ptr(rv), len(rh), _ptr(ah), _ptr(av), len(ah))
        ids = np.zeros(max(batch.num_insertions, 1), np.int64)
        s = _lib.HyperSummary()
        _lib.check(_lib.lib().dyng_hip_hypergraph_apply(self._h, C.byref(nb), int(max_weight), C.byref(s),
                                                        ids.ctypes.data))
        return HyperApplySummary(**{f: int(getattr(s, f)) for f in _lib.HYPER_SUMMARY_FIELDS},
                                 inserted_ids=ids[:batch.num_insertions].copy())
