"""Weighted hypergraphs with slack_csr storage: the container dyng_amd.hyper_sssp repairs its
shortest hyperpaths on (python/dyng/hypergraph.py, cpp/include/dyng/hypergraph/hypergraph.hpp).

What the path needs of the reference's container, with its names and rules:

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
- the ids the insertions get: the store's best-fit id allocator (id_allocator.cpp:119-580,
  restated in :class:`_IdAllocator`), under the batch order of ``id_reuse`` (``erase_first``:
  deletions, removals, additions, insertions; ``insert_first``: insertions first) and the
  payload sizing, compaction and growth of the ``CapacityPolicy`` (capacity.cpp:105-170,
  store_policy.cpp:71-116).

This is host bookkeeping (Python); the line graph's shortest paths run on the device
(dyng_amd.hyper_sssp). The escher storage, unweighted hypergraphs' algorithms and the
hypergraph file formats are not built.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Iterable, NamedTuple

import numpy as np

from .errors import DyngError, InvalidArgumentError, NotSupportedError

__all__ = ["CapacityError", "CapacityPolicy", "HypergraphProperties", "HyperApplySummary", "HyperedgeBatch",
           "Hypergraph", "Incidence", "LineGraph"]

INT_MAX = 2**31 - 1
MAX_WEIGHT = 2**28


class CapacityError(DyngError):
    """The storage cannot hold the batch (error.hpp's capacity_error)."""


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


def _padded4(count: int) -> int:
    """Padding rules 1 and 3: the values, zeros, one terminator slot (id_allocator.cpp:39-41)."""
    return ((count + 4) // 4) * 4


def _aligned(count: int, alignment: int) -> int:
    """Padding rule 2: a surplus row's slots for its values (id_allocator.cpp:45-47)."""
    return 0 if count <= 0 else (count + alignment - 1) // alignment * alignment


_ABSENT, _LIVE, _DEAD = 0, 1, 2


class _IdAllocator:
    """The best-fit id allocator of the hyperedge tree (id_allocator.cpp:119-580): per key
    (id + 1) its state, slot length, value count and tail segment; the bump pointer and the
    payload capacity. Only what decides the ids is kept: the values themselves only for rows
    whose tail is an overflow segment (a later removal leaves in the tail's occupancy what is
    left there)."""

    def __init__(self, policy: CapacityPolicy, counts: list[int]):
        self.grows = policy.grow_factor != 0.0
        self.compacts = policy.compact_before_grow
        self.alignment = policy.row_alignment
        self.headroom = policy.compaction_headroom
        self.max_payload = policy.limit() if self.grows else INT_MAX
        size = sum(_padded4(c) for c in counts)
        self.capacity = policy.initial_payload(size)
        n = len(counts)
        self.state = [_ABSENT] + [_LIVE] * n
        self.length = [0] + [_padded4(c) for c in counts]
        self.count = [0] + list(counts)
        self.tail_capacity = [0] + [_padded4(c) - 1 for c in counts]
        self.occupancy = [0] + list(counts)
        self.tails: dict[int, list[int]] = {}
        self.dead_keys: set[int] = set()
        self.max_key = n
        self.live = n
        self.dead = 0
        self.bump = size
        self.compactions = 0
        self._log: list | None = None  # (key, state, length, count, tail_capacity, occupancy, tail) to undo

    def _at(self, key: int) -> None:
        while len(self.state) <= key:
            self.state.append(_ABSENT)
            self.length.append(0)
            self.count.append(0)
            self.tail_capacity.append(0)
            self.occupancy.append(0)

    def _touch(self, key: int) -> None:
        """Record a key's slot before it changes (undone by rollback())."""
        if self._log is not None and key < len(self.state):
            self._log.append((key, self.state[key], self.length[key], self.count[key], self.tail_capacity[key],
                              self.occupancy[key], list(self.tails[key]) if key in self.tails else None))

    def begin(self) -> None:
        """Start recording: rollback() restores the state of this moment (a failed apply)."""
        self._log = []
        self._saved = (len(self.state), set(self.dead_keys), self.max_key, self.live, self.dead, self.bump,
                       self.capacity, self.compactions)

    def commit(self) -> None:
        self._log = None

    def rollback(self) -> None:
        log, self._log = self._log, None
        for key, st, ln, ct, tc, oc, tail in reversed(log or []):
            self.state[key], self.length[key], self.count[key] = st, ln, ct
            self.tail_capacity[key], self.occupancy[key] = tc, oc
            if tail is None:
                self.tails.pop(key, None)
            else:
                self.tails[key] = tail
        size, dead_keys, self.max_key, self.live, self.dead, self.bump, self.capacity, self.compactions = self._saved
        self.dead_keys = dead_keys
        for lst in (self.state, self.length, self.count, self.tail_capacity, self.occupancy):
            del lst[size:]

    def _clear(self, key: int) -> None:
        self._touch(key)
        self.state[key] = _ABSENT
        self.length[key] = self.count[key] = self.tail_capacity[key] = self.occupancy[key] = 0
        self.tails.pop(key, None)
        self.dead_keys.discard(key)

    def _grown(self, needed: int) -> int:
        """store_policy::grown_capacity (store_policy.cpp:100-116)."""
        if not self.grows:
            raise CapacityError(f"dyng: CBST h2v: {needed} payload entries do not fit its capacity of "
                                f"{self.capacity} (the policy does not grow the payload)")
        grown = min(max(needed + needed // 4, 3 * self.capacity // 2), self.max_payload)
        if needed > grown:
            raise CapacityError(f"dyng: CBST h2v: {needed} payload entries exceed the largest payload "
                                f"(at most {self.max_payload})")
        return grown

    def _ensure(self, needed: int) -> None:
        if needed > self.capacity:
            self.capacity = self._grown(needed)

    def _keep_headroom(self, compactions_before: int) -> None:
        """The hysteresis after an operation that compacted (id_allocator.cpp:212-225)."""
        if self.headroom <= 0 or self.compactions == compactions_before or not self.grows:
            return
        slack = self.capacity // self.headroom
        if self.capacity - self.bump >= slack:
            return
        self._ensure(min(self.bump + slack, min(self.max_payload, INT_MAX)))

    def compact(self) -> None:
        """Every row one segment in key order (id_allocator.cpp:227-263)."""
        if self.live + self.dead == 0:
            return
        total = 0
        for key in range(1, self.max_key + 1):
            if self.state[key] == _LIVE:
                total += _padded4(self.count[key])
            elif self.state[key] == _DEAD:
                total += self.length[key]
        if total > INT_MAX:
            raise CapacityError(f"dyng: CBST h2v: the compacted rows need {total} payload slots")
        self._ensure(total)
        for key in range(1, self.max_key + 1):
            self._touch(key)
            if self.state[key] == _LIVE:
                self.length[key] = _padded4(self.count[key])
                self.tail_capacity[key] = self.length[key] - 1
                self.occupancy[key] = self.count[key]
            elif self.state[key] == _DEAD:
                self.tail_capacity[key] = self.length[key] - 1
                self.occupancy[key] = 0
        self.bump = total
        self.compactions += 1
        self.tails.clear()

    def erase(self, keys: list[int]) -> None:
        """Mark live keys as deleted slots (id_allocator.cpp:265-282)."""
        for key in keys:
            if 0 < key < len(self.state) and self.state[key] == _LIVE:
                self._touch(key)
                self.state[key] = _DEAD
                self.dead_keys.add(key)
                self.live -= 1
                self.dead += 1
                self.tails.pop(key, None)

    def insert(self, counts: list[int]) -> list[int]:
        """Best-fit reuse of the deleted slots, then the surplus rows after the last live key
        (id_allocator.cpp:284-411)."""
        num = len(counts)
        keys = [0] * num
        if num == 0:
            return keys
        compactions_before = self.compactions
        matched = [False] * num
        reuse_k = min(num, self.dead)
        if reuse_k > 0:
            dead_keys = sorted(self.dead_keys)
            cap = [self.length[k] - 1 for k in dead_keys]
            slot_order = sorted(range(len(dead_keys)), key=lambda j: cap[j])  # stable: key order on ties
            sorted_cap = [cap[j] for j in slot_order]
            item_order = sorted(range(reuse_k), key=lambda i: counts[i])  # stable: input order on ties
            prefix_max = None
            for i, item in enumerate(item_order):
                lo = _lower_bound(sorted_cap, counts[item])
                prefix_max = lo - i if prefix_max is None else max(prefix_max, lo - i)
                assigned = i + prefix_max
                if assigned >= len(dead_keys):
                    break
                key = dead_keys[slot_order[assigned]]
                self._touch(key)
                self.state[key] = _LIVE
                self.dead_keys.discard(key)
                self.count[key] = counts[item]
                self.tail_capacity[key] = self.length[key] - 1
                self.occupancy[key] = counts[item]
                self.dead -= 1
                self.live += 1
                keys[item] = key
                matched[item] = True
        surplus = [i for i in range(num) if not matched[i]]
        if surplus:
            slots = sum(_aligned(counts[i], self.alignment) + 1 for i in surplus)
            if self.bump + slots > self.capacity and self.compacts:
                self.compact()
            self._ensure(self.bump + slots)
            self.bump += slots
            # the rebuild: deleted slots not reused are dropped; the surplus rows get the keys after
            # the last live key, in item order
            for key in list(self.dead_keys):
                self._clear(key)
            self.dead = 0
            last_live = self.max_key
            while last_live > 0 and self.state[last_live] != _LIVE:
                last_live -= 1
            nxt = last_live + 1
            for i in surplus:
                a = _aligned(counts[i], self.alignment)
                self._at(nxt)
                self._touch(nxt)
                self.state[nxt] = _LIVE
                self.length[nxt] = a + 1
                self.count[nxt] = counts[i]
                self.tail_capacity[nxt] = a
                self.occupancy[nxt] = counts[i]
                self.tails.pop(nxt, None)
                self.live += 1
                keys[i] = nxt
                nxt += 1
            for key in range(nxt, self.max_key + 1):
                self._clear(key)
            self.max_key = nxt - 1
        self._keep_headroom(compactions_before)
        return keys

    def append(self, groups: list[tuple[int, list[int]]]) -> None:
        """Append values to live rows, key by key (a fill, id_allocator.cpp:413-498)."""
        before = self.compactions
        self._fill(groups, allow_compact=True)
        self._keep_headroom(before)

    def _fill(self, groups: list[tuple[int, list[int]]], allow_compact: bool) -> None:
        if not groups:
            return
        taken, segment = [], []
        total = 0
        for key, values in groups:
            self._touch(key)
            occ = min(max(self.occupancy[key], 0), self.tail_capacity[key])
            room = self.tail_capacity[key] - occ
            k = len(values)
            fit = min(k, room)
            if key in self.tails:
                self.tails[key].extend(values[:fit])
            if k <= room:
                self.occupancy[key] = occ + k
                self.count[key] += k
                taken.append(k)
                segment.append(0)
            else:
                self.occupancy[key] = self.tail_capacity[key]
                self.count[key] += room
                taken.append(room)
                seg = _padded4(k - room)
                segment.append(seg)
                total += seg
        needed = self.bump + total
        if total > 0 and needed > self.capacity and allow_compact and self.compacts:
            rest = [(key, values[taken[i]:]) for i, (key, values) in enumerate(groups) if segment[i] != 0]
            self.compact()
            self._fill(rest, allow_compact=False)
            return
        if total > 0:
            self._ensure(needed)
            for i, (key, values) in enumerate(groups):
                if segment[i] == 0:
                    continue
                self.count[key] += len(values) - taken[i]
                self.tail_capacity[key] = segment[i] - 1
                self.occupancy[key] = len(values) - taken[i]
                self.tails[key] = list(values[taken[i]:])
            self.bump = needed

    def remove_values(self, groups: list[tuple[int, list[int]]]) -> None:
        """Remove values from live rows (an unfill, id_allocator.cpp:500-540)."""
        for key, values in groups:
            if not (0 < key < len(self.state)) or self.state[key] != _LIVE:
                continue
            self._touch(key)
            removed = set(values)
            self.count[key] -= len(removed)
            if key not in self.tails:
                self.occupancy[key] = self.count[key]
            else:
                tail = [v for v in self.tails[key] if v not in removed]
                self.tails[key] = tail
                self.occupancy[key] = len(tail)

    def copy(self) -> "_IdAllocator":
        out = object.__new__(_IdAllocator)
        out.__dict__.update({k: (list(v) if isinstance(v, list) else v) for k, v in self.__dict__.items()})
        out.tails = {k: list(v) for k, v in self.tails.items()}
        out.dead_keys = set(self.dead_keys)
        out._log = None
        return out


def _lower_bound(a: list[int], x: int) -> int:
    lo, hi = 0, len(a)
    while lo < hi:
        mid = (lo + hi) // 2
        if a[mid] < x:
            lo = mid + 1
        else:
            hi = mid
    return lo


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


@dataclass(frozen=True)
class _Applied:
    """What an apply changed, for the results that follow it: the ids whose rows changed and
    their rows before and after (empty: dead), and the summary."""

    summary: HyperApplySummary
    pre: dict[int, list[int]]
    post: dict[int, list[int]]
    weights_pre: dict[int, int]


class Hypergraph:
    """A weighted hypergraph with slack_csr storage (hypergraph.hpp; python/dyng/hypergraph.py)."""

    def __init__(self) -> None:
        raise TypeError("use dyng_amd.Hypergraph.from_lists()")

    @classmethod
    def from_lists(cls, lists: Iterable[Iterable[int]], weights: Iterable[int] | None = None, *,
                   num_vertices: int | None = None, storage: str = "slack_csr", line_graph: str = "automatic",
                   id_reuse: str = "erase_first", capacity: CapacityPolicy | str = "dyng_default",
                   store_v2h: bool = True, resources: Any = None) -> "Hypergraph":
        """Hyperedge i gets id i; repeated members are dropped (hypergraph.cpp, from_lists)."""
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
        raw = [[int(v) for v in h] for h in lists]
        rows = [sorted(set(h)) for h in raw]
        n = int(num_vertices) if num_vertices is not None else (max((max(r) for r in rows if r), default=-1) + 1)
        for r in rows:
            if r and (r[0] < 0 or r[-1] >= n):
                raise InvalidArgumentError(f"dyng_amd.Hypergraph.from_lists: a member is outside [0, {n})")
        w = None
        if weights is not None:
            w = [int(x) for x in weights]
            if len(w) != len(rows):
                raise InvalidArgumentError(f"dyng_amd.Hypergraph.from_lists: {len(w)} weights for {len(rows)} hyperedges")
            if any(x < 1 for x in w):
                raise InvalidArgumentError("dyng_amd.Hypergraph.from_lists: weights are >= 1")
        self = object.__new__(cls)
        self._n = n
        self._rows = rows
        self._w = w
        self._version = 0
        self._props = HypergraphProperties("slack_csr", line_graph, id_reuse, policy, True)
        self._alloc = _IdAllocator(policy, [len(r) for r in rows])
        self._v2h: list[set[int]] = [set() for _ in range(n)]
        for h, r in enumerate(rows):
            for v in r:
                self._v2h[v].add(h)
        self._live = sum(1 for r in rows if r)
        return self

    # ---- views ----

    @property
    def num_vertices(self) -> int:
        return self._n

    @property
    def num_hyperedges(self) -> int:
        return self._live

    @property
    def hyperedge_id_bound(self) -> int:
        return self._alloc.max_key

    @property
    def version(self) -> int:
        return self._version

    @property
    def properties(self) -> HypergraphProperties:
        return self._props

    @property
    def weighted(self) -> bool:
        return self._w is not None

    @property
    def space(self) -> str:
        return "host"

    @property
    def weights(self) -> np.ndarray | None:
        """The weight of every id below the bound (0 for a dead id)."""
        if self._w is None:
            return None
        b = self.hyperedge_id_bound
        return np.asarray([self._w[h] if h < len(self._rows) and self._rows[h] else 0 for h in range(b)], np.int64)

    def rows(self) -> list[list[int]]:
        """The members of every id below the bound (empty for a dead id)."""
        b = self.hyperedge_id_bound
        return [list(self._rows[h]) if h < len(self._rows) else [] for h in range(b)]

    def h2v(self) -> Incidence:
        return _csr(self.rows(), self._version, Incidence)

    def v2h(self) -> Incidence:
        return _csr([sorted(s) for s in self._v2h], self._version, Incidence)

    def neighbours(self, h: int) -> set[int]:
        """The live ids sharing a vertex with id h."""
        if h >= len(self._rows) or not self._rows[h]:
            return set()
        out: set[int] = set()
        for v in self._rows[h]:
            out |= self._v2h[v]
        out.discard(h)
        return out

    def h2h(self) -> LineGraph:
        """The line graph (ordered pairs of live ids that share a vertex, rows sorted)."""
        return _csr([sorted(self.neighbours(h)) for h in range(self.hyperedge_id_bound)], self._version, LineGraph)

    def clone(self) -> "Hypergraph":
        out = object.__new__(Hypergraph)
        out._n = self._n
        out._rows = [list(r) for r in self._rows]
        out._w = None if self._w is None else list(self._w)
        out._version = self._version
        out._props = self._props
        out._alloc = self._alloc.copy()
        out._v2h = [set(s) for s in self._v2h]
        out._live = self._live
        return out

    def __copy__(self) -> "Hypergraph":
        return self.clone()

    def __deepcopy__(self, memo: dict) -> "Hypergraph":
        return self.clone()

    def __repr__(self) -> str:
        return (f"dyng_amd.Hypergraph(num_vertices={self._n}, num_hyperedges={self.num_hyperedges}, "
                f"hyperedge_id_bound={self.hyperedge_id_bound}, version={self._version})")

    # ---- apply ----

    def apply(self, batch: HyperedgeBatch) -> HyperApplySummary:
        """Apply a batch (hypergraph::apply); the version grows by one."""
        return self._apply(batch).summary

    def _check(self, batch: HyperedgeBatch, validate_weights: bool = False) -> list[list[int]]:
        """The checks before anything changes; the inserted rows (members sorted, deduplicated)."""
        if not isinstance(batch, HyperedgeBatch):
            raise TypeError("dyng_amd.Hypergraph.apply: expected a dyng_amd.HyperedgeBatch")
        rows = []
        for h in batch._insertions:
            if not h:
                raise InvalidArgumentError("dyng_amd.Hypergraph.apply: an inserted hyperedge has no member")
            if min(h) < 0 or max(h) >= self._n:
                raise InvalidArgumentError(f"dyng_amd.Hypergraph.apply: a member of an inserted hyperedge is "
                                           f"outside [0, {self._n})")
            rows.append(sorted(set(h)))
        if self.weighted:
            if batch.num_insertions and batch._weights is None:
                raise InvalidArgumentError("dyng_amd.Hypergraph.apply: a weighted hypergraph needs one insert weight "
                                           "per inserted hyperedge")
            if batch._weights is not None and len(batch._weights) and int(batch._weights.min()) < 1:
                raise InvalidArgumentError("dyng_amd.Hypergraph.apply: insert weights are >= 1")
            if validate_weights and batch._weights is not None and len(batch._weights) \
                    and int(batch._weights.max()) > MAX_WEIGHT:
                raise InvalidArgumentError(f"dyng_amd.hyper_sssp.update: an inserted weight is outside [1, 2^28]")
        return rows

    def _apply(self, batch: HyperedgeBatch, validate_weights: bool = False) -> _Applied:
        rows_in = self._check(batch, validate_weights)
        bound = self.hyperedge_id_bound

        def live(h: int) -> bool:
            return 0 <= h < bound and h < len(self._rows) and bool(self._rows[h])

        # deletions: checked against the hypergraph before the batch (hypergraph.cpp:855-889)
        deleted: list[int] = []
        seen: set[int] = set()
        ignored = 0
        for h in batch._delete.tolist():
            if live(h) and h not in seen:
                seen.add(h)
                deleted.append(h)
            else:
                ignored += 1
        dropped = sum(len(h) for h in batch._insertions) - sum(len(r) for r in rows_in)
        # incidence edits (hypergraph.cpp:613-760): per hyperedge in id order, its removals then its
        # additions in batch order, each against the row as the earlier edits left it
        rem_h, rem_v = batch._del_inc
        add_h, add_v = batch._ins_inc
        nrem = len(rem_h)
        edits: dict[int, list[tuple[int, int]]] = {}
        for j in range(nrem + len(add_h)):
            h = int(rem_h[j]) if j < nrem else int(add_h[j - nrem])
            v = int(rem_v[j]) if j < nrem else int(add_v[j - nrem])
            if 0 <= h < bound and 0 <= v < self._n:
                edits.setdefault(h, []).append((j, v))
        removals: list[tuple[int, list[int]]] = []
        additions: list[tuple[int, list[int]]] = []
        edited_rows: dict[int, list[int]] = {}
        applied_edits = removed = added = 0
        for h in sorted(edits):
            if not live(h) or h in seen:
                continue
            row = list(self._rows[h])
            rem, add = [], []
            for j, v in edits[h]:
                member = v in row
                if j < nrem:
                    if not member or len(row) == 1:
                        continue
                    row.remove(v)
                    rem.append(v)
                    removed += 1
                else:
                    if member:
                        continue
                    row.append(v)
                    row.sort()
                    add.append(v)
                    added += 1
            if rem or add:
                edited_rows[h] = row
            if rem:
                removals.append((h + 1, rem))
            if add:
                additions.append((h + 1, add))
        applied_edits = removed + added
        skipped_edits = nrem + len(add_h) - applied_edits
        # the store phase (hypergraph.cpp:945-1075): the allocator in the order of id_reuse
        alloc = self._alloc
        counts = [len(r) for r in rows_in]
        alloc.begin()  # a step that fails (CapacityError) leaves the allocator as it was
        try:
            if self._props.reuse == "insert_first":
                keys = alloc.insert(counts)
                alloc.erase([h + 1 for h in deleted])
                alloc.remove_values(removals)
                alloc.append(additions)
            else:
                alloc.erase([h + 1 for h in deleted])
                alloc.remove_values(removals)
                alloc.append(additions)
                keys = alloc.insert(counts)
        except BaseException:
            alloc.rollback()
            raise
        alloc.commit()
        ids = [k - 1 for k in keys]
        # commit the rows
        pre: dict[int, list[int]] = {}
        post: dict[int, list[int]] = {}
        weights_pre: dict[int, int] = {}

        def row_of(h: int) -> list[int]:
            return list(self._rows[h]) if h < len(self._rows) else []

        for h in deleted:
            pre.setdefault(h, row_of(h))
        for h in edited_rows:
            pre.setdefault(h, row_of(h))
        for h in ids:
            pre.setdefault(h, row_of(h))
        for h in pre:
            if self._w is not None and h < len(self._w):
                weights_pre[h] = self._w[h]
        for h in deleted:
            self._set_row(h, [])
        for h, row in edited_rows.items():
            self._set_row(h, row)
        for k, h in enumerate(ids):
            self._set_row(h, rows_in[k])
            if self._w is not None:
                while len(self._w) <= h:
                    self._w.append(0)
                self._w[h] = int(batch._weights[k])
        for h in pre:
            post[h] = row_of(h)
        self._version += 1
        summary = HyperApplySummary(
            inserted_hyperedges=len(rows_in), deleted_hyperedges=len(deleted), inserted_incidences=added,
            deleted_incidences=removed, skipped_ops=ignored + dropped + skipped_edits, ignored_deletions=ignored,
            dropped_members=dropped, num_hyperedges_after=self.num_hyperedges, id_bound_after=self.hyperedge_id_bound,
            inserted_ids=np.asarray(ids, np.int64))
        return _Applied(summary, pre, post, weights_pre)

    def _set_row(self, h: int, row: list[int]) -> None:
        while len(self._rows) <= h:
            self._rows.append([])
        self._live += (1 if row else 0) - (1 if self._rows[h] else 0)
        for v in self._rows[h]:
            self._v2h[v].discard(h)
        self._rows[h] = list(row)
        for v in row:
            self._v2h[v].add(h)
