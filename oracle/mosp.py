"""CPU restatement of MOSP's combined solve (test infrastructure only: the checker of the HIP
path's `dyng_hip_mosp_combined`; nothing in the product imports it).

Follows the reference's host path (paths relative to the reference root):

- ``preference_scale``: L = lcm(preferences), 1 for none, at most 2^20
  (cpp/src/algorithms/mosp/mosp.cpp:114-131);
- ``combine``: the combined graph of the K trees: for every vertex v other than the source and
  every objective k with parent p = parent_k[v] >= 0, the edge (p, v) once (its first objective
  k), weighing base - sum of terms[j] over the objectives j whose tree holds it, with
  base = L (K + 1) and terms[k] = L / pref[k] (L without preferences)
  (cpp/src/algorithms/mosp/openmp.cpp:39-118, mosp.cpp:283-289);
- the solve: the canonical shortest-path tree of the combined graph from the source (the
  oracle's Dijkstra with lowest-id tight parents; sssp_solve_view, cpp/src/algorithms/sssp/sssp.cpp:1251-1286);
- ``path_costs``: the objective values of the MOSP path to every vertex along the combined tree,
  each tree edge (p, v) taken as the first (p, v) of p's row of the graph
  (cpp/src/algorithms/mosp/sequential.cpp:172-236); unreachable vertices keep INF.
- ``combine_fast`` / ``path_costs_fast``: the same two functions over whole arrays (numpy), for
  graphs where the loops above take too long (the CPU baseline); tests/test_mosp_oracle.py checks
  them against the loops.

Pinned by the reference's committed MOSP combined fixtures (cpp/tests/data/mosp_combined:
distances, trees and path costs of 8 preference cases; tests/test_oracle.py) and the thesis
worked example (cost (15, 3, 20) at vertex 6, python/dyng/mosp.py docstring).
"""

from __future__ import annotations

import math

import numpy as np

from . import INF, Csr, dijkstra

MAX_PREFERENCE_SCALE = 1 << 20


def preference_scale(preferences, num_objectives: int) -> int:
    if preferences is None or len(preferences) == 0:
        return 1
    if len(preferences) != num_objectives:
        raise ValueError(f"mosp: {len(preferences)} preferences for {num_objectives} objectives")
    scale = 1
    for p in preferences:
        if p < 1:
            raise ValueError("mosp: a preference must be >= 1")
        scale = scale // math.gcd(scale, int(p)) * int(p)
        if scale > MAX_PREFERENCE_SCALE:
            raise ValueError("mosp: lcm(preferences) exceeds 2^20")
    return scale


def terms_and_base(preferences, num_objectives: int) -> tuple[list[int], int, int]:
    scale = preference_scale(preferences, num_objectives)
    if preferences is None or len(preferences) == 0:
        terms = [scale] * num_objectives
    else:
        terms = [scale // int(p) for p in preferences]
    return terms, scale * (num_objectives + 1), scale


def combine(parents, source: int, preferences=None) -> tuple[Csr, int, int]:
    """The combined graph (rows sorted by head id) of the K parent arrays; also its weight sum
    and the base weight."""
    k_count = len(parents)
    parents = [np.asarray(p, np.int64) for p in parents]
    n = len(parents[0])
    terms, base, _ = terms_and_base(preferences, k_count)
    src, dst, w = [], [], []
    for v in range(n):
        if v == source:
            continue
        for k in range(k_count):
            p = int(parents[k][v])
            if p < 0 or any(int(parents[j][v]) == p for j in range(k)):
                continue
            weight = base - terms[k] - sum(terms[j] for j in range(k + 1, k_count) if int(parents[j][v]) == p)
            src.append(p)
            dst.append(v)
            w.append(weight)
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    w = np.asarray(w, np.int64)
    order = np.lexsort((dst, src))
    src, dst, w = src[order], dst[order], w[order]
    row = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(src, minlength=n), out=row[1:])
    g = Csr(row, dst.astype(np.int32), w.astype(np.int32).reshape(1, -1))
    return g, int(w.sum()), base


def solve(parents, source: int, preferences=None) -> tuple[np.ndarray, np.ndarray, int]:
    """The MOSP tree: (combined distances, combined parents, combined edge count)."""
    g, _, _ = combine(parents, source, preferences)
    d, p = dijkstra(g, source, 0)
    return d, p, g.m


def path_costs(g: Csr, parent, source: int, columns: int) -> np.ndarray:
    """(n, columns) int64: the first `columns` weights summed along the tree path from the source;
    INF where the tree does not reach. Raises if a tree edge is not an edge of g."""
    parent = np.asarray(parent, np.int64)
    n = g.n
    costs = np.full((n, columns), INF, np.int64)
    if n == 0:
        return costs
    children = [[] for _ in range(n)]
    for v in range(n):
        if v != source and parent[v] >= 0:
            children[int(parent[v])].append(v)
    costs[source, :] = 0
    queue = [source]
    i = 0
    while i < len(queue):
        p = queue[i]
        i += 1
        lo, hi = int(g.row_ptr[p]), int(g.row_ptr[p + 1])
        for v in children[p]:
            hit = np.nonzero(g.col[lo:hi] == v)[0]
            if len(hit) == 0:
                raise ValueError(f"mosp: the MOSP tree edge ({p}, {v}) is not an edge of the graph")
            e = lo + int(hit[0])
            costs[v, :] = costs[p, :] + g.w[:columns, e].astype(np.int64)
            queue.append(v)
    return costs


def combine_fast(parents, source: int, preferences=None) -> tuple[Csr, int, int]:
    """``combine`` over whole arrays: the same graph, weight sum and base."""
    k_count = len(parents)
    P = np.stack([np.asarray(p, np.int64) for p in parents])
    n = P.shape[1]
    terms, base, _ = terms_and_base(preferences, k_count)
    v = np.arange(n, dtype=np.int64)
    src, dst, w = [], [], []
    for k in range(k_count):
        p = P[k]
        keep = (p >= 0) & (v != source)
        for j in range(k):
            keep &= P[j] != p
        weight = np.full(n, base - terms[k], np.int64)
        for j in range(k + 1, k_count):
            weight -= np.where(P[j] == p, terms[j], 0)
        src.append(p[keep])
        dst.append(v[keep])
        w.append(weight[keep])
    src = np.concatenate(src) if src else np.zeros(0, np.int64)
    dst = np.concatenate(dst) if dst else np.zeros(0, np.int64)
    w = np.concatenate(w) if w else np.zeros(0, np.int64)
    order = np.lexsort((dst, src))
    src, dst, w = src[order], dst[order], w[order]
    row = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(src, minlength=n), out=row[1:])
    g = Csr(row, dst.astype(np.int32), w.astype(np.int32).reshape(1, -1))
    return g, int(w.sum()), base


def path_costs_fast(g: Csr, parent, source: int, columns: int) -> np.ndarray:
    """``path_costs`` over whole arrays: each tree edge (p, v) located as the first (p, v) of p's
    row by a stable sort of the (row, neighbour) keys, then path doubling up the tree."""
    parent = np.asarray(parent, np.int64)
    n = g.n
    costs = np.full((n, columns), INF, np.int64)
    if n == 0:
        return costs
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(g.row_ptr))
    keys = rows * n + g.col.astype(np.int64)
    if len(keys) < 2 or bool((keys[1:] >= keys[:-1]).all()):
        order = np.arange(len(keys))  # rows already sorted by neighbour
        sorted_keys = keys
    else:
        order = np.argsort(keys, kind="stable")
        sorted_keys = keys[order]
    v = np.arange(n, dtype=np.int64)
    has = (parent >= 0) & (v != source)
    want = parent[has] * n + v[has]
    at = np.searchsorted(sorted_keys, want, side="left")
    found = at < len(sorted_keys)
    found[found] = sorted_keys[at[found]] == want[found]
    if not found.all():
        bad = v[has][~found][0]
        raise ValueError(f"mosp: the MOSP tree edge ({int(parent[bad])}, {int(bad)}) is not an edge of the graph")
    acc = np.zeros((n, columns), np.int64)
    acc[v[has]] = g.w[:columns, order[at]].T.astype(np.int64)
    # acc[v] sums the edges from v up to end[v]; each round doubles the covered segments
    end = np.where(has, parent, v)
    rounds = 2
    while (1 << rounds) < n:
        rounds += 1
    for _ in range(rounds + 1):
        active = has[end]
        if not active.any():
            break
        e = end[active]
        acc[active] += acc[e]
        nxt = end.copy()
        nxt[active] = end[e]
        end = nxt
    # the vertices whose path runs back to the source hold costs (a parent cycle never ends)
    reach = (end == source) & ~has[end]
    costs[reach] = acc[reach]
    return costs
