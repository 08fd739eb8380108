"""Helpers of the GPU tests that call the engine seam of include/dyng_hip.h directly: a graph and
the tree of its source, one batch applied to the graph alone (dyng_hip_graph_apply), and a
`dyng_hip_sssp_run` filled the way a `backend::hip` branch of `sssp_problem::resume`
(sssp.cpp:691-823) would fill `sssp_run` (problem.hpp:380-400) before calling the engine.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

import oracle


@dataclass
class SeamCase:
    graph: object        # dyng_amd.Graph, after the batch
    old_dist: np.ndarray  # the tree before the batch (the engine's own, canonical)
    old_parent: np.ndarray
    want: oracle.Tree     # the oracle's tree after the batch
    counts: oracle.Counters
    max_weight: int
    delta: int


def make_case(dg, n: int, m: int, ops: int, seed: int, wmax: int = 100) -> SeamCase:
    """A random directed graph without self-loops (w ~ U[1, wmax]), the tree of source 0, and one
    mixed batch (ops // 2 new insertions, ops // 2 deletions of existing edges) applied to the graph
    alone. The oracle updates its own copy of the tree."""
    rng = np.random.default_rng(seed)
    src, dst = rng.integers(0, n, m), rng.integers(0, n, m)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    w = rng.integers(1, wmax + 1, len(src))
    og = oracle.build_from_edges(n, src, dst, w.reshape(-1, 1))
    g = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    ot, _ = oracle.compute(og, 0)
    r = dg.sssp.compute(g, 0)
    d0, p0 = r.to_numpy()
    assert np.array_equal(d0, ot.dist) and np.array_equal(p0, ot.parent), "compute differs from the oracle"
    rows = np.repeat(np.arange(og.n), np.diff(og.row_ptr))
    k = ops // 2
    pick = rng.choice(og.m, k, replace=False)
    dels = (rows[pick], og.col[pick].copy())
    ins = (rng.integers(0, n, k), rng.integers(0, n, k), rng.integers(1, wmax + 1, k))
    ins = ((ins[0] + (ins[0] == ins[1])) % n, ins[1], ins[2])  # no self-loops
    _, _, counts = oracle.update(og, ot, ins[0], ins[1], ins[2].reshape(-1, 1), dels[0], dels[1])
    g.apply(dg.EdgeBatch(insert=ins, delete=dels))
    return SeamCase(g, d0, p0, ot, counts, wmax, 32)


def fill_run(dg, case: SeamCase, dist, par, workspace=None, check_cycles: int = 0):
    """The engine run of `case`: the graph after the batch, the old tree in (dist, par) (device
    arrays, overwritten by the new tree), the objective's largest weight and the near-far width."""
    from dyng_amd import _lib

    run = _lib.SsspRun()
    run.graph = case.graph._handle.value
    run.objective = 0
    run.source = 0
    run.delta = case.delta
    run.max_weight = case.max_weight
    run.distances = dist.ptr
    run.parents = par.ptr
    run.check_cycles = check_cycles
    run.workspace = workspace
    run.tuning = _lib.Tuning()
    run.sources = None
    run.num_sources = 0
    run.tree_token = 0
    return run


def run_update(dg, case: SeamCase, entry, workspace=None, check_cycles: int = 0):
    """Call an update entry point of the seam on a device copy of the old tree; returns the new
    tree (host copies) and the counters."""
    from dyng_amd import _lib

    dist, par = dg.DeviceArray(case.old_dist), dg.DeviceArray(case.old_parent)
    run = fill_run(dg, case, dist, par, workspace, check_cycles)
    counters = _lib.Counters()
    _lib.check(entry(C.byref(run), None, C.byref(counters)))
    return dist.to_numpy(), par.to_numpy(), counters


def workspace(dg):
    """A device workspace of the engine seam (dyng_hip_workspace_create), destroyed with the
    returned guard."""
    from dyng_amd import _lib

    h = C.c_void_p()
    _lib.check(_lib.lib().dyng_hip_workspace_create(0, C.byref(h)))

    class _Guard:
        handle = h.value

        def __enter__(self):
            return self.handle

        def __exit__(self, *exc):
            _lib.lib().dyng_hip_workspace_destroy(self.handle)

    return _Guard()
