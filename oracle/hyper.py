"""CPU restatement of dynG's hyper_sssp checker (test infrastructure only: the oracle of the HIP
path's ``dyng_amd.hyper_sssp``; nothing in the product imports it).

Follows the reference's test oracle (paths relative to the reference root), the definition the
reference's own hyper_sssp tests check every backend against:

- ``line_graph``: every ordered pair (a, b), a != b, of live hyperedges that share a vertex,
  sorted and deduplicated; a hyperedge is live when its member row is not empty
  (cpp/src/testing/hyper_sssp_reference.cpp:60-119, rebuild);
- ``reference``: from a source vertex s, every live hyperedge h that holds s is seeded with
  D(h) = w(h); Dijkstra where stepping into a neighbour b costs w(b); the canonical parent of h
  is its lowest-id live neighbour p with D(p) + w(h) == D(h), and -1 for a seeded hyperedge with
  D(h) == w(h), a dead or an unreachable one (hyper_sssp_reference.cpp:133-220, solve);
- ``line_graph_delta``: the unordered pairs {a, b} (a < b) a batch removed from and added to
  the line graph (hyper_sssp_reference.cpp:248-289);
- ``vertex_distances``: a vertex's distance is the smallest D(h) over its live hyperedges, 0 for
  the source (python/dyng/hyper_sssp.py, Result.distance_to).

Pinned in tests/test_hyper_oracle.py by the reference's worked example and test expectations
(python/dyng/hyper_sssp.py docstring; python/tests/test_hyper_sssp.py:63-82,104-112), by
Bellman-Ford over the line graph on small random hypergraphs, and by planted mistakes.
"""

from __future__ import annotations

import heapq

INF = (2**63 - 1) // 4


def line_graph(rows: list[list[int]]) -> list[list[int]]:
    """The neighbours of every id (sorted): the live ids sharing a vertex with it."""
    by_vertex: dict[int, list[int]] = {}
    for h, row in enumerate(rows):
        for v in set(row):
            by_vertex.setdefault(v, []).append(h)
    out = [set() for _ in rows]
    for hs in by_vertex.values():
        for a in hs:
            for b in hs:
                if a != b:
                    out[a].add(b)
    return [sorted(s) for s in out]


def reference(rows: list[list[int]], weights: list[int], source: int) -> tuple[list[int], list[int]]:
    """(distances, parents) by id from the source vertex."""
    m = len(rows)
    lg = line_graph(rows)
    dist = [INF] * m
    parent = [-1] * m
    seeded = [False] * m
    heap: list[tuple[int, int]] = []
    for h, row in enumerate(rows):
        if source in row:
            seeded[h] = True
            if weights[h] < dist[h]:
                dist[h] = weights[h]
                heapq.heappush(heap, (dist[h], h))
    while heap:
        d, u = heapq.heappop(heap)
        if d != dist[u]:
            continue
        for v in lg[u]:
            nd = d + weights[v]
            if nd < dist[v]:
                dist[v] = nd
                heapq.heappush(heap, (nd, v))
    for h in range(m):
        if dist[h] >= INF or (seeded[h] and dist[h] == weights[h]):
            continue
        for p in lg[h]:
            if dist[p] < INF and dist[p] + weights[h] == dist[h]:
                parent[h] = p
                break
    return dist, parent


def line_graph_delta(pre: list[list[int]], post: list[list[int]]) -> tuple[list[tuple[int, int]], list[tuple[int, int]]]:
    """(deleted, inserted): the unordered pairs (a < b) the batch removed and added."""
    a = line_graph(pre)
    b = line_graph(post)
    m = max(len(a), len(b))
    deleted, inserted = [], []
    for i in range(m):
        r0 = set(a[i]) if i < len(a) else set()
        r1 = set(b[i]) if i < len(b) else set()
        deleted += [(i, j) for j in sorted(r0 - r1) if j > i]
        inserted += [(i, j) for j in sorted(r1 - r0) if j > i]
    return deleted, inserted


def vertex_distances(rows: list[list[int]], dist: list[int], num_vertices: int, source: int) -> list[int]:
    """The distance of every vertex: min D(h) over its live hyperedges, 0 for the source."""
    out = [INF] * num_vertices
    for h, row in enumerate(rows):
        for v in set(row):
            out[v] = min(out[v], dist[h])
    if 0 <= source < num_vertices:
        out[source] = 0
    return out
