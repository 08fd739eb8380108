"""Synthetic graphs and batches of the benchmark configurations (libdyng_gen.so, host OpenMP).

Every draw is splitmix64(seed, draw index): inputs are identical for every run, thread count and
leg (GPU path, CPU oracle). Graph seed 12345 and batch seed 777, each xor the configuration index.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdyng_gen.so")

GRAPH_SEED = 12345
BATCH_SEED = 777
RMAT_ABC = (0.57, 0.19, 0.19)

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build

            build.build_generators()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.dyng_gen_splitmix.restype = C.c_uint64
        L.dyng_gen_splitmix.argtypes = [C.c_uint64, C.c_uint64]
        L.dyng_gen_er.restype = C.c_int64
        L.dyng_gen_er.argtypes = [C.c_int64, C.c_int64, C.c_int32, C.c_uint64, vp, vp, vp]
        L.dyng_gen_rmat.restype = C.c_int64
        L.dyng_gen_rmat.argtypes = [C.c_int32, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_int32,
                                    C.c_uint64, vp, vp, vp]
        L.dyng_gen_grid.restype = C.c_int64
        L.dyng_gen_grid.argtypes = [C.c_int64, C.c_double, C.c_int32, C.c_uint64, vp, vp, vp, C.c_int64]
        L.dyng_gen_batches.restype = C.c_int
        L.dyng_gen_batches.argtypes = [vp, vp, C.c_int64, C.c_int64, C.c_int32, C.c_int64, C.c_int64, C.c_int64,
                                       C.c_int32, C.c_uint64, C.c_int32, C.c_double, C.c_double, C.c_double,
                                       vp, vp, vp, vp, vp]
        L.dyng_gen_stream.restype = C.c_int64
        L.dyng_gen_stream.argtypes = [vp, vp, C.c_int64, C.c_int64, C.c_int32, C.c_double, C.c_double, C.c_int32,
                                      C.c_uint64, C.c_int32, C.c_double, C.c_double, C.c_double, vp, vp, vp, vp,
                                      vp, vp]
        L.dyng_gen_tree_batch.restype = C.c_int
        L.dyng_gen_tree_batch.argtypes = [C.c_int64, vp, C.c_int64, C.c_int64, C.c_int64, C.c_int32, C.c_uint64,
                                          C.c_int64, vp, vp, vp, vp, vp]
        _lib = L
    return _lib


def splitmix(seed: int, i: int) -> int:
    return int(lib().dyng_gen_splitmix(seed & (2**64 - 1), i))


def _out(m: int):
    return np.empty(m, np.int32), np.empty(m, np.int32), np.empty(m, np.int32)


def er(n: int, m: int, wmax: int, seed: int):
    """Directed Erdos-Renyi: m distinct non-loop edges with uniform endpoints, w ~ U[1, wmax]."""
    s, d, w = _out(m)
    got = lib().dyng_gen_er(n, m, wmax, seed, s.ctypes.data, d.ctypes.data, w.ctypes.data)
    if got != m:
        raise ValueError("er: cannot draw that many distinct edges")
    return s, d, w


def rmat(scale: int, m: int, wmax: int, seed: int, abc=RMAT_ABC):
    """RMAT (a, b, c, d = 1 - a - b - c): m distinct non-loop edges, w ~ U[1, wmax]."""
    s, d, w = _out(m)
    got = lib().dyng_gen_rmat(scale, m, abc[0], abc[1], abc[2], wmax, seed, s.ctypes.data, d.ctypes.data,
                              w.ctypes.data)
    if got != m:
        raise ValueError("rmat: cannot draw that many distinct edges")
    return s, d, w


def grid(side: int, shortcut_fraction: float, wmax: int, seed: int):
    """side x side grid, 4-neighbour arcs both ways, plus ceil(fraction * arcs) random shortcuts."""
    arcs = 4 * side * (side - 1)
    cap = arcs + int(np.ceil(shortcut_fraction * arcs)) + 1
    s, d, w = _out(cap)
    got = lib().dyng_gen_grid(side, shortcut_fraction, wmax, seed, s.ctypes.data, d.ctypes.data, w.ctypes.data,
                              cap)
    if got < 0:
        raise ValueError("grid: capacity")
    return s[:got].copy(), d[:got].copy(), w[:got].copy()


def mixed_batches(src: np.ndarray, dst: np.ndarray, n: int, batches: int, n_ins: int, n_del: int, n_rew: int,
                  wmax: int, seed: int, rmat_scale: int = 0, abc=RMAT_ABC):
    """Batches against the graph (src, dst sorted and distinct): new-edge insertions, deletions of
    existing edges and reweights (upserts of existing edges); see generators.cpp."""
    src = np.ascontiguousarray(src, dtype=np.int32)
    dst = np.ascontiguousarray(dst, dtype=np.int32)
    per_ins = n_ins + n_rew
    ins_s, ins_d, ins_w = (np.empty(batches * per_ins, np.int32) for _ in range(3))
    del_s, del_d = (np.empty(batches * n_del, np.int32) for _ in range(2))
    rc = lib().dyng_gen_batches(src.ctypes.data, dst.ctypes.data, len(src), n, batches, n_ins, n_del, n_rew, wmax,
                                seed, rmat_scale, abc[0], abc[1], abc[2], ins_s.ctypes.data, ins_d.ctypes.data,
                                ins_w.ctypes.data, del_s.ctypes.data, del_d.ctypes.data)
    if rc != 0:
        raise ValueError("mixed_batches: the graph has too few edges for the requested batches")
    out = []
    for b in range(batches):
        i0, i1 = b * per_ins, (b + 1) * per_ins
        d0, d1 = b * n_del, (b + 1) * n_del
        out.append((ins_s[i0:i1], ins_d[i0:i1], ins_w[i0:i1], del_s[d0:d1], del_d[d0:d1]))
    return out


def max_out_degree_vertex(src: np.ndarray, n: int) -> int:
    """The vertex with the largest out-degree, lowest id on ties."""
    deg = np.bincount(src, minlength=n)
    return int(np.argmax(deg))


def tree_batch(side: int, parents: np.ndarray, n_ins: int, n_tree: int, n_rand: int, wmax: int, seed: int,
               batch: int):
    """A batch against the current tree of a side x side grid graph: n_tree deletions of distinct
    tree edges (parent[v], v), n_rand deletions of random grid arcs, n_ins random insertions with
    w ~ U[1, wmax] (generators.cpp: dyng_gen_tree_batch)."""
    parents = np.ascontiguousarray(parents, dtype=np.int32)
    if len(parents) != side * side:
        raise ValueError("tree_batch: parents must have side * side entries")
    ins_s, ins_d, ins_w = (np.empty(n_ins, np.int32) for _ in range(3))
    del_s, del_d = (np.empty(n_tree + n_rand, np.int32) for _ in range(2))
    rc = lib().dyng_gen_tree_batch(side, parents.ctypes.data, n_ins, n_tree, n_rand, wmax, seed, batch,
                                   ins_s.ctypes.data, ins_d.ctypes.data, ins_w.ctypes.data, del_s.ctypes.data,
                                   del_d.ctypes.data)
    if rc != 0:
        raise ValueError("tree_batch: the tree has too few edges")
    return ins_s, ins_d, ins_w, del_s, del_d


def stream(src: np.ndarray, dst: np.ndarray, n: int, count: int, lo: float, hi: float, wmax: int, seed: int,
           rmat_scale: int, abc=RMAT_ABC):
    """Batches of log-uniform size in [lo, hi] with a uniform delete fraction against the graph
    (src, dst sorted and distinct): deletions of distinct existing edges, insertions of new RMAT
    pairs (generators.cpp: dyng_gen_stream). Returns a list of (ins_src, ins_dst, ins_w, del_src,
    del_dst)."""
    src = np.ascontiguousarray(src, dtype=np.int32)
    dst = np.ascontiguousarray(dst, dtype=np.int32)
    sizes = np.zeros(2 * count, np.int64)
    L = lib()
    L.dyng_gen_stream(src.ctypes.data, dst.ctypes.data, len(src), n, count, lo, hi, wmax, seed, rmat_scale,
                      abc[0], abc[1], abc[2], sizes.ctypes.data, None, None, None, None, None)
    ni, nd = int(sizes[0::2].sum()), int(sizes[1::2].sum())
    ins_s, ins_d, ins_w = (np.empty(max(ni, 1), np.int32) for _ in range(3))
    del_s, del_d = (np.empty(max(nd, 1), np.int32) for _ in range(2))
    L.dyng_gen_stream(src.ctypes.data, dst.ctypes.data, len(src), n, count, lo, hi, wmax, seed, rmat_scale,
                      abc[0], abc[1], abc[2], sizes.ctypes.data, ins_s.ctypes.data, ins_d.ctypes.data,
                      ins_w.ctypes.data, del_s.ctypes.data, del_d.ctypes.data)
    out, i0, d0 = [], 0, 0
    for b in range(count):
        a, d = int(sizes[2 * b]), int(sizes[2 * b + 1])
        out.append((ins_s[i0:i0 + a], ins_d[i0:i0 + a], ins_w[i0:i0 + a], del_s[d0:d0 + d], del_d[d0:d0 + d]))
        i0 += a
        d0 += d
    return out
