"""Readers and writers of MOSP's text formats, for the golden fixtures under tests/golden.

Formats (cpp/include/dyng/io/csr_triplet.hpp, batch_io.hpp:582-623, result_io.hpp:25-80):
RowPtr/ColInd hold one integer per line; Values one line per edge with K weights; insert.txt
"u v w1 .. wK"; delete.txt "u v"; distance files "v d" or "v INF"; tree files "v p".
"""

from __future__ import annotations

import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
INF = (2**63 - 1) // 4


def _ints(path: str) -> list[list[int]]:
    rows = []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if parts:
                rows.append([int(x) for x in parts])
    return rows


def read_csr(prefix: str, num_weights: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    row = np.array([r[0] for r in _ints(prefix + "RowPtr.txt")], np.int64)
    col = np.array([r[0] for r in _ints(prefix + "ColInd.txt")], np.int32)
    vals = _ints(prefix + "Values.txt")
    w = np.array(vals, np.int64).reshape(len(col), num_weights) if len(col) else \
        np.zeros((0, num_weights), np.int64)
    return row, col, np.ascontiguousarray(w.T.astype(np.int32))


def read_batch(directory: str, num_weights: int):
    ins = _ints(os.path.join(directory, "insert.txt"))
    dele = _ints(os.path.join(directory, "delete.txt"))
    ins_src = np.array([r[0] for r in ins], np.int32)
    ins_dst = np.array([r[1] for r in ins], np.int32)
    ins_w = np.array([r[2:2 + num_weights] for r in ins], np.int64).reshape(len(ins), num_weights)
    del_src = np.array([r[0] for r in dele], np.int32)
    del_dst = np.array([r[1] for r in dele], np.int32)
    return ins_src, ins_dst, ins_w, del_src, del_dst


def read_distances(path: str, n: int) -> np.ndarray:
    d = np.full(n, INF, np.int64)
    with open(path) as f:
        for line in f:
            parts = line.split()
            if parts:
                d[int(parts[0])] = INF if parts[1] == "INF" else int(parts[1])
    return d


def read_parents(path: str, n: int) -> np.ndarray:
    p = np.full(n, -1, np.int32)
    with open(path) as f:
        for line in f:
            parts = line.split()
            if parts:
                p[int(parts[0])] = int(parts[1])
    return p


def distances_text(d) -> str:
    """write_distances (result_io.cpp:165-180): 'v d', 'v INF' at or above INF / 2."""
    return "".join(f"{v} {'INF' if x >= INF // 2 else int(x)}\n" for v, x in enumerate(d))


def parents_text(p) -> str:
    return "".join(f"{v} {int(x)}\n" for v, x in enumerate(p))


def read_text(path: str) -> str:
    with open(path) as f:
        return f.read()


def sssp_cases() -> list[tuple[str, str, int]]:
    """(name, input directory, K) of mosp_sssp/cases.txt."""
    out = []
    for r in open(os.path.join(GOLDEN, "mosp_sssp", "cases.txt")):
        parts = r.split()
        if parts:
            out.append((parts[0], os.path.join(GOLDEN, parts[1]), int(parts[2])))
    return out


def expected_invalidated(name: str, k: int) -> int:
    for line in open(os.path.join(GOLDEN, "mosp_sssp", name, "stats.txt")):
        parts = line.split()
        if parts and parts[0] == f"obj{k}":
            return int(parts[2])
    return -1


def apply_cases() -> list[str]:
    base = os.path.join(GOLDEN, "mosp_graph_io")
    return [c for c in sorted(os.listdir(base))
            if os.path.isdir(os.path.join(base, c, "applied"))]


def combined_cases() -> list[tuple[str, str, str, int, list[int]]]:
    """(name, input directory, init directory, K, preferences) of mosp_combined/cases.txt."""
    out = []
    for r in open(os.path.join(GOLDEN, "mosp_combined", "cases.txt")):
        parts = r.split()
        if parts:
            out.append((parts[0], os.path.join(GOLDEN, parts[1]), os.path.join(GOLDEN, parts[2]), int(parts[3]),
                        [int(x) for x in parts[4].split(",")]))
    return out


def num_weight_columns(prefix: str) -> int:
    """Weights per line of <prefix>Values.txt."""
    with open(prefix + "Values.txt") as f:
        for line in f:
            if line.split():
                return len(line.split())
    return 1


def read_costs(path: str, n: int) -> np.ndarray:
    """mospCosts.txt: 'v c1 .. cK' per vertex, INF for unreachable."""
    rows = {}
    width = 0
    with open(path) as f:
        for line in f:
            parts = line.split()
            if parts:
                rows[int(parts[0])] = [INF if x == "INF" else int(x) for x in parts[1:]]
                width = len(parts) - 1
    c = np.full((n, width), INF, np.int64)
    for v, r in rows.items():
        c[v] = r
    return c
