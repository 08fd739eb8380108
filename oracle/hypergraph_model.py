"""CPU restatement of the rules of dynG's hypergraph container that decide results and ids
(test infrastructure only: the checker of the C++ container behind dyng_amd.Hypergraph; nothing
in the product imports it).

Follows the reference (paths relative to its root):

- the batch phases and skip rules: hyperedge deletions (a negative, out-of-range, dead or repeated
  id is skipped), incidence removals, incidence additions (per hyperedge in id order, its removals
  then its additions in batch order, each against the row the earlier edits left; an edit of a
  hyperedge that is not live or that the batch deletes, of a vertex out of range, a removal of a
  non-member or of the last member, an addition of a member, is skipped), hyperedge insertions
  (repeated members dropped) (cpp/src/hypergraph/hypergraph.cpp:613-760, 855-889,
  cpp/include/dyng/hypergraph/hyperedge_batch.hpp:27-40);
- the order of the store phase under id_reuse (hypergraph.cpp:945-1075);
- the best-fit id allocator: slot lengths (padding rules 1-3), best fit of the first min(K, D)
  items into the deleted slots, the surplus rows after the last live key, compaction, growth and
  the headroom after a compaction (cpp/src/hypergraph/id_allocator.cpp:39-47,119-540), with the
  payload sizing and growth of the capacity policy (cpp/src/hypergraph/capacity.cpp:105-170,
  cpp/src/hypergraph/escher/store_policy.cpp:71-116).

Pinned in tests/test_hyper_oracle.py by the reference's own expectations (its id-reuse example,
skipped-operation and incidence-edit counts, python/tests/test_hypergraph.py:50-70,195-258) and
by the reference test model's validity rules on random batches
(python/dyng/testing/oracles.py:129-235).
"""

from __future__ import annotations

from dataclasses import dataclass

INT_MAX = 2**31 - 1
MAX_WEIGHT = 2**28


class CapacityError(Exception):
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
            raise ValueError(f"CapacityPolicy.preset: unknown preset {name!r}")
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


class Model:
    """A weighted hypergraph's rows, weights and allocator, batch by batch."""

    def __init__(self, lists, weights, num_vertices: int, reuse: str = "erase_first",
                 policy: CapacityPolicy | None = None):
        self.rows = [sorted(set(int(v) for v in h)) for h in lists]
        self.w = None if weights is None else [int(x) for x in weights]
        self.n = int(num_vertices)
        self.reuse = reuse
        self.alloc = _IdAllocator(policy or CapacityPolicy(), [len(r) for r in self.rows])

    @property
    def id_bound(self) -> int:
        return self.alloc.max_key

    def rows_by_id(self) -> list[list[int]]:
        return [list(self.rows[h]) if h < len(self.rows) else [] for h in range(self.id_bound)]

    def apply(self, dels, ins, weights, rem=((), ()), add=((), ())):
        """Apply a batch; returns (summary counts, inserted ids). Raises CapacityError (nothing
        changed) when the storage cannot hold it."""
        bound = self.id_bound
        rows_in = [sorted(set(int(v) for v in h)) for h in ins]

        def live(h):
            return 0 <= h < bound and h < len(self.rows) and bool(self.rows[h])

        deleted, seen, ignored = [], set(), 0
        for h in dels:
            h = int(h)
            if live(h) and h not in seen:
                seen.add(h)
                deleted.append(h)
            else:
                ignored += 1
        dropped = sum(len(h) for h in ins) - sum(len(r) for r in rows_in)
        rem_h, rem_v = [list(x) for x in rem]
        add_h, add_v = [list(x) for x in add]
        nrem = len(rem_h)
        edits = {}
        for j in range(nrem + len(add_h)):
            h = int(rem_h[j]) if j < nrem else int(add_h[j - nrem])
            v = int(rem_v[j]) if j < nrem else int(add_v[j - nrem])
            if 0 <= h < bound and 0 <= v < self.n:
                edits.setdefault(h, []).append((j, v))
        removals, additions, edited = [], [], {}
        removed = added = 0
        for h in sorted(edits):
            if not live(h) or h in seen:
                continue
            row = list(self.rows[h])
            r_, a_ = [], []
            for j, v in edits[h]:
                member = v in row
                if j < nrem:
                    if not member or len(row) == 1:
                        continue
                    row.remove(v)
                    r_.append(v)
                    removed += 1
                else:
                    if member:
                        continue
                    row.append(v)
                    row.sort()
                    a_.append(v)
                    added += 1
            if r_ or a_:
                edited[h] = row
            if r_:
                removals.append((h + 1, r_))
            if a_:
                additions.append((h + 1, a_))
        skipped = nrem + len(add_h) - removed - added
        alloc = self.alloc.copy()
        counts = [len(r) for r in rows_in]
        if self.reuse == "insert_first":
            keys = alloc.insert(counts)
            alloc.erase([h + 1 for h in deleted])
            alloc.remove_values(removals)
            alloc.append(additions)
        else:
            alloc.erase([h + 1 for h in deleted])
            alloc.remove_values(removals)
            alloc.append(additions)
            keys = alloc.insert(counts)
        self.alloc = alloc
        ids = [k - 1 for k in keys]
        for h in deleted:
            self._set(h, [])
        for h, row in edited.items():
            self._set(h, row)
        for k, h in enumerate(ids):
            self._set(h, rows_in[k])
            if self.w is not None:
                while len(self.w) <= h:
                    self.w.append(0)
                self.w[h] = int(weights[k])
        counts_out = {"inserted_hyperedges": len(rows_in), "deleted_hyperedges": len(deleted),
                      "inserted_incidences": added, "deleted_incidences": removed,
                      "skipped_ops": ignored + dropped + skipped, "ignored_deletions": ignored,
                      "dropped_members": dropped, "num_hyperedges_after": sum(1 for r in self.rows if r),
                      "id_bound_after": self.id_bound}
        return counts_out, ids

    def _set(self, h, row):
        while len(self.rows) <= h:
            self.rows.append([])
        self.rows[h] = list(row)
