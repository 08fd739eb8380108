"""Graph shapes and brute-force checks shared by the CPU and GPU tests.

The grids and the path follow the reference's distance-only and packing-boundary tests
(cpp/tests/algorithms/sssp/sssp_test.cpp:239-315, sssp_cuda_test.cpp:501-617). `brute_tree` is an
independent restatement of what a canonical shortest-path tree is: distances from scipy's
Dijkstra (float64, exact below 2^53) and, for every reachable vertex but the source, the lowest-id
in-neighbour u with d(u) + w(u, v) == d(v) (check_sssp.hpp:56-85), found with numpy, not with the
oracle's C code.
"""

from __future__ import annotations

import numpy as np

INF = (1 << 61) - 1  # sssp_infinity: numeric_limits<int64>::max() / 4 (types.hpp:91)


def grid_one_way(side: int, w: int):
    """Arcs right and down only (sssp_test.cpp:239-252)."""
    v = np.arange(side * side, dtype=np.int64).reshape(side, side)
    right = (v[:, :-1].ravel(), v[:, 1:].ravel())
    down = (v[:-1, :].ravel(), v[1:, :].ravel())
    src = np.concatenate([right[0], down[0]])
    dst = np.concatenate([right[1], down[1]])
    return side * side, src, dst, np.full(len(src), w, np.int64)


def grid_two_way(side: int):
    """Arcs both ways between 4-neighbours (sssp_test.cpp:282-299)."""
    n, s, d, _ = grid_one_way(side, 1)
    return n, np.concatenate([s, d]), np.concatenate([d, s])


def path(n: int, w: int, back_edge: bool):
    """0 -> 1 -> ... -> n-1, each of weight w, plus n-1 -> 1 when back_edge (sssp_cuda_test.cpp:506-520)."""
    src = np.arange(n - 1, dtype=np.int64)
    dst = src + 1
    if back_edge:
        src = np.append(src, n - 1)
        dst = np.append(dst, 1)
    return src, dst, np.full(len(src), w, np.int64)


def brute_tree(n: int, src, dst, w, source: int):
    """The canonical tree of the graph (src, dst, w): (int64 distances with INF, int32 parents)."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import dijkstra

    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    w = np.asarray(w, np.int64)
    # parallel edges: scipy would add them up, keep the lightest of each pair
    order = np.lexsort((w, dst, src))
    s, t, x = src[order], dst[order], w[order]
    first = np.ones(len(s), bool)
    first[1:] = (s[1:] != s[:-1]) | (t[1:] != t[:-1])
    s, t, x = s[first], t[first], x[first]
    m = csr_matrix((x.astype(np.float64), (s, t)), shape=(n, n))
    df = dijkstra(m, directed=True, indices=source)
    reach = np.isfinite(df)
    assert not reach.any() or df[reach].max() < 2.0**53, "distances beyond float64's exact integers"
    dist = np.full(n, INF, np.int64)
    dist[reach] = df[reach].astype(np.int64)
    ok = reach[s] & reach[t] & (t != source)
    tight = ok & (dist[np.where(ok, s, 0)] + x == dist[t])
    parent = np.full(n, np.iinfo(np.int32).max, np.int64)
    np.minimum.at(parent, t[tight], s[tight])
    parent[~reach] = -1
    parent[source] = -1
    assert (parent[reach] != np.iinfo(np.int32).max).all() or reach.sum() <= 1
    return dist, parent.astype(np.int32)


def subtree_union(parent, roots) -> int:
    """Size of the union of the subtrees of `parent` rooted at `roots` (numpy, iterative)."""
    parent = np.asarray(parent, np.int64)
    n = len(parent)
    mark = np.zeros(n, bool)
    roots = np.asarray(sorted(set(int(r) for r in roots)), np.int64)
    if len(roots) == 0:
        return 0
    mark[roots] = True
    has = parent >= 0
    while True:
        grow = has & ~mark
        grow[grow] = mark[parent[grow]]
        if not grow.any():
            return int(mark.sum())
        mark |= grow


def random_tight_parents(rng, n, src, dst, w, dist, parent, source):
    """A valid shortest-path tree that is not canonical: every vertex with several tight
    in-neighbours takes one of them at random (acyclic, since weights are >= 1)."""
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    w = np.asarray(w, np.int64)
    reach = dist < INF // 2
    ok = reach[src] & reach[dst] & (dst != source)
    tight = np.where(ok & (dist[np.where(ok, src, 0)] + w == dist[dst]))[0]
    out = parent.astype(np.int32).copy()
    key = rng.random(len(tight))
    order = np.lexsort((key, dst[tight]))
    pick = tight[order]
    first = np.ones(len(pick), bool)
    first[1:] = dst[pick][1:] != dst[pick][:-1]
    pick = pick[first]
    out[dst[pick]] = src[pick].astype(np.int32)
    return out
