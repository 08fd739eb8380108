"""Many independent shortest-path trees repaired together (the batched multi-instance path; it has
no counterpart in the reference, which updates one result per call).

Independent graphs are laid out as one device graph, their disjoint union: instance i's vertex v
is the union's vertex offsets[i] + v. One ``sssp.compute_many`` run computes every instance's
tree (one source each) in one launch, and one ``sssp.update`` applies all the instances' batches
(one device apply) and repairs all the trees in one launch. Each instance's result is exactly
what a run on the instance alone gives: the union has no edge between instances, every offset
keeps the order of ids (the lowest-id tie rule), and the near-far width only changes the
schedule.

``autotune`` picks the schedule knobs (row-length threshold, chunk size, near-far width) of a
repair by timing candidates on clones of a result against one applied batch.
"""

from __future__ import annotations

import dataclasses
from typing import Sequence

import numpy as np

from . import sssp
from .batch import EdgeBatch
from .errors import InvalidArgumentError
from .graph import Graph, GraphProperties


@dataclasses.dataclass(frozen=True)
class InstanceBatch:
    """One instance's batch in its own vertex ids."""

    ins_src: np.ndarray
    ins_dst: np.ndarray
    ins_w: np.ndarray
    del_src: np.ndarray
    del_dst: np.ndarray


class Instances:
    """Independent graphs held as one device graph, repaired together.

    ``graphs`` is a sequence of (num_vertices, src, dst, weights) in each instance's own ids.
    """

    def __init__(self, graphs: Sequence[tuple], properties: GraphProperties | None = None, resources=None):
        if not graphs:
            raise InvalidArgumentError("batched.Instances: no graphs")
        sizes = [int(g[0]) for g in graphs]
        if min(sizes) < 1:
            raise InvalidArgumentError("batched.Instances: every instance needs a vertex")
        self.sizes = np.asarray(sizes, np.int64)
        self.offsets = np.zeros(len(sizes) + 1, np.int64)
        np.cumsum(self.sizes, out=self.offsets[1:])
        if self.offsets[-1] >= 2**31 - 1:
            raise InvalidArgumentError("batched.Instances: the union exceeds int32 vertex ids")
        src, dst, w = [], [], []
        for i, (n, s, d, wt) in enumerate(graphs):
            s = np.asarray(s, np.int64)
            d = np.asarray(d, np.int64)
            if s.size and (s.min() < 0 or s.max() >= n or d.min() < 0 or d.max() >= n):
                raise InvalidArgumentError(f"batched.Instances: instance {i} names a vertex outside [0, {n})")
            src.append(s + self.offsets[i])
            dst.append(d + self.offsets[i])
            w.append(np.asarray(wt).reshape(len(s), -1) if len(s) else np.zeros((0, 1), np.int64))
        self.graph = Graph.from_edges(np.concatenate(src), np.concatenate(dst), np.concatenate(w),
                                      num_vertices=int(self.offsets[-1]), properties=properties,
                                      resources=resources)

    @property
    def count(self) -> int:
        return len(self.sizes)

    def compute(self, sources: Sequence[int], **kwargs) -> sssp.Result:
        """Every instance's tree from its source, in one launch."""
        src = np.asarray(sources, np.int64)
        if len(src) != self.count:
            raise InvalidArgumentError(f"batched.Instances.compute: {len(src)} sources for {self.count} instances")
        if (src < 0).any() or (src >= self.sizes).any():
            raise InvalidArgumentError("batched.Instances.compute: a source is outside its instance")
        return sssp.compute_many(self.graph, src + self.offsets[:-1], **kwargs)

    def batch(self, batches: Sequence[InstanceBatch | tuple]) -> EdgeBatch:
        """The union's batch: every instance's operations in union ids, instance by instance (the
        instances share no edge, so their order does not change any instance's result)."""
        if len(batches) != self.count:
            raise InvalidArgumentError(f"batched.Instances.batch: {len(batches)} batches for {self.count} instances")
        parts = [[], [], [], [], []]
        for i, b in enumerate(batches):
            b = b if isinstance(b, InstanceBatch) else InstanceBatch(*b)
            off, n = self.offsets[i], self.sizes[i]
            for a, name in ((b.ins_src, "insertion"), (b.ins_dst, "insertion"), (b.del_src, "deletion"),
                            (b.del_dst, "deletion")):
                a = np.asarray(a, np.int64)
                if a.size and (a.min() < 0 or a.max() >= n):
                    raise InvalidArgumentError(f"batched.Instances.batch: instance {i} has a {name} outside "
                                               f"[0, {n}) (instances cannot grow)")
            parts[0].append(np.asarray(b.ins_src, np.int64) + off)
            parts[1].append(np.asarray(b.ins_dst, np.int64) + off)
            parts[2].append(np.asarray(b.ins_w, np.int64).reshape(len(b.ins_src), -1))
            parts[3].append(np.asarray(b.del_src, np.int64) + off)
            parts[4].append(np.asarray(b.del_dst, np.int64) + off)
        cat = [np.concatenate(p) for p in parts]
        return EdgeBatch(insert=(cat[0], cat[1], cat[2]), delete=(cat[3], cat[4]))

    def update(self, batches: Sequence[InstanceBatch | tuple], result: sssp.Result) -> sssp.Stats:
        """Apply every instance's batch and repair every tree: one device apply, one launch."""
        return sssp.update(self.graph, self.batch(batches), result)

    def tree(self, result: sssp.Result, i: int, arrays=None) -> tuple[np.ndarray, np.ndarray]:
        """Instance i's (distances, parents) in its own ids (``arrays``: the union's, if already
        downloaded)."""
        d, p = arrays if arrays is not None else result.to_numpy()
        lo, hi = int(self.offsets[i]), int(self.offsets[i + 1])
        di = d[lo:hi].copy()
        pi = p[lo:hi].copy()
        pi[pi >= 0] -= lo
        return di, pi


def default_delta(graph: Graph, objective: int = 0) -> int:
    """sssp_default_delta (sssp.cpp:136-145): max(1, 32 * average weight / average degree)."""
    _, weight_sum = graph.column(objective)
    m, n = max(graph.num_edges, 1), max(graph.num_vertices, 1)
    return max(1, int(32.0 * (weight_sum / m) / (m / n)))


def candidate_tunings(delta: int) -> list[tuple[sssp.Tuning, int]]:
    """The schedule grid the autotuner times: row-length thresholds for the edge-balanced tiles,
    chunk sizes for the long rows and near-far widths around the default."""
    out = []
    for light in (256, 1024, 4096):
        for chunk in (128, 512):
            for scale in (0.25, 1.0, 4.0):
                out.append((sssp.Tuning(light_degree=light, chunk_edges=chunk), max(1, int(delta * scale))))
    return out


@dataclasses.dataclass
class TuneReport:
    tuning: sssp.Tuning
    delta: int
    engine_ms: float
    table: list  # (tuning, delta, engine_ms) per candidate


def autotune(graph: Graph, batch: EdgeBatch, result: sssp.Result, candidates=None, repeats: int = 1) -> TuneReport:
    """Apply ``batch`` to ``graph``, time the repair of clones of ``result`` under each candidate
    schedule, then repair ``result`` itself with the fastest one (its tuning and delta stay set on
    it). Every candidate gives the same tree; only the time differs."""
    if candidates is None:
        candidates = candidate_tunings(default_delta(graph, result.options.objective))
    graph.apply(batch)
    table = []
    for tuning, delta in candidates:
        best = None
        for _ in range(max(1, repeats)):
            trial = result.clone()
            trial.set_tuning(tuning)
            trial.set_options(delta=delta)
            st = sssp.repair(graph, trial)
            best = st.device.kernel_ms if best is None else min(best, st.device.kernel_ms)
            del trial
        table.append((tuning, delta, best))
    tuning, delta, ms = min(table, key=lambda t: t[2])
    result.set_tuning(tuning)
    result.set_options(delta=delta)
    sssp.repair(graph, result)
    return TuneReport(tuning, delta, ms, table)
