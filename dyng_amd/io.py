"""dyng_amd.io: MOSP's text files (python/dyng/io.py; cpp/include/dyng/io/csr_triplet.hpp,
batch_io.hpp, result_io.hpp).

- the text CSR ``<prefix>RowPtr.txt`` / ``ColInd.txt`` / ``Values.txt`` (MOSP's readCsrGraph /
  writeCsrGraph; cpp/src/io/csr_triplet.cpp);
- the batch files ``insert.txt`` (``u v w1 .. wK``) and ``delete.txt`` (``u v``) (MOSP's
  readChangeBatch / writeChangeBatch; cpp/src/io/batch_io.cpp:27-139), with dynG's strict rules
  or, with ``mosp_lenient``, MOSP's: integers after the first 2 + K of an insertion line and the
  first 2 of a deletion line are ignored, and a deletion line with one integer is skipped;
- the result files: distances ``v d`` / ``v INF``, trees ``v p``, path costs ``v c1 .. cK``
  (MOSP's writeDistances / writeParents / the ``mosp`` driver's writeCosts;
  cpp/src/io/result_io.cpp:109-260). A distance or cost at or above INFINITE_DISTANCE / 2 is
  written ``INF``; ``INF`` reads as INFINITE_DISTANCE.

Readers are strict: a malformed file raises :class:`FileFormatError` naming the file and line.
Writers produce the bytes MOSP's own writers produce. These are host-side files; the graph and the
trees live on the device, so reading a graph builds the device graph and writing a tree copies it
to the host first.
"""

from __future__ import annotations

import os
from pathlib import Path
from typing import Any

import numpy as np

from .batch import EdgeBatch
from .errors import DyngError, InvalidArgumentError
from .graph import Graph, GraphProperties
from .resources import Resources

__all__ = ["FileFormatError", "read_csr_triplet", "write_csr_triplet", "read_legacy_batch", "write_legacy_batch",
           "read_distances", "read_parents", "write_distances", "write_parents", "write_path_costs"]

INFINITE_DISTANCE = (2**63 - 1) // 4
_INT32_MAX = 2**31 - 1


class FileFormatError(DyngError, ValueError):
    """A file is missing or malformed (io_error, cpp/include/dyng/core/error.hpp); ``path`` and
    ``line`` (1-based, 0 when the whole file is at fault) name the place."""

    def __init__(self, path: str, line: int, message: str):
        where = f"{path}:{line}" if line else path
        super().__init__(f"dyng: {where}: {message}")
        self.path = path
        self.line = line


def _text(path) -> str:
    p = os.fspath(path)
    try:
        with open(p, "rb") as f:
            data = f.read()
    except OSError as e:
        raise FileFormatError(p, 0, f"cannot read the file ({e.strerror})") from None
    try:
        return data.decode("ascii")
    except UnicodeDecodeError as e:
        raise FileFormatError(p, data.count(b"\n", 0, e.start) + 1, "not ASCII text") from None


def _int(path: str, line: int, token: str, lo: int, hi: int, what: str) -> int:
    try:
        value = int(token, 10)
    except ValueError:
        raise FileFormatError(path, line, f"not an integer {what}: {token!r}") from None
    if not token.lstrip("+-").isdigit():
        raise FileFormatError(path, line, f"not an integer {what}: {token!r}")
    if value < lo or value > hi:
        raise FileFormatError(path, line, f"{what} {value} is out of range [{lo}, {hi}]")
    return value


def _lines(text: str):
    """(1-based line number, its tokens) of every non-blank line."""
    for number, line in enumerate(text.split("\n"), 1):
        fields = line.split()
        if fields:
            yield number, fields


def _numbers(path: str, text: str, lo: int, hi: int, what: str) -> np.ndarray:
    """Every whitespace-separated integer of the file, in [lo, hi] (a fast pass; the line of the
    first bad token is found by a second, slow one)."""
    try:
        a = np.array(text.split(), dtype=np.int64)
        if a.size == 0 or (a.min() >= lo and a.max() <= hi):
            return a
    except (ValueError, OverflowError):
        pass
    for number, fields in _lines(text):
        for tok in fields:
            _int(path, number, tok, lo, hi, what)
    raise FileFormatError(path, 0, f"a malformed {what}")  # not reached


# ------------------------------------------------------------------------------------------------
# The text CSR
# ------------------------------------------------------------------------------------------------


def _csr_files(prefix) -> tuple[str, str, str]:
    p = os.fspath(prefix)
    return p + "RowPtr.txt", p + "ColInd.txt", p + "Values.txt"


def read_csr_arrays(prefix, *, num_weights: int = 0) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(row_ptr int64 (n+1), col int32 (m), weights int32 (m, K)) of MOSP's text CSR; K from the
    first line of Values.txt (``num_weights`` for a graph without edges)."""
    rp_path, ci_path, va_path = _csr_files(prefix)
    row = _numbers(rp_path, _text(rp_path), 0, 2**62, "row offset")
    if len(row) < 2:
        raise FileFormatError(rp_path, 0, "a CSR needs at least 2 row offsets (one vertex)")
    if row[0] != 0:
        raise FileFormatError(rp_path, 1, "the first row offset must be 0")
    if (np.diff(row) < 0).any():
        i = int(np.nonzero(np.diff(row) < 0)[0][0]) + 1
        raise FileFormatError(rp_path, i + 1, f"row offsets must not decrease ({int(row[i])} after "
                                              f"{int(row[i - 1])})")
    n, m = len(row) - 1, int(row[-1])
    if n > _INT32_MAX:
        raise FileFormatError(rp_path, 0, "too many vertices for int32 ids")
    col = _numbers(ci_path, _text(ci_path), 0, n - 1, "column index")
    if len(col) != m:
        raise FileFormatError(ci_path, 0, f"{len(col)} column indices, but the row offsets announce {m} edges")
    vtext = _text(va_path)
    k = int(num_weights)
    rows = []
    for number, fields in _lines(vtext):
        if k == 0:
            k = len(fields)
        elif len(fields) != k:
            raise FileFormatError(va_path, number, f"{len(fields)} weights on the line, expected {k}")
        rows.append(number)
    if len(rows) != m:
        raise FileFormatError(va_path, 0, f"{len(rows)} weight lines, but the row offsets announce {m} edges")
    if k == 0:
        k = 1
    w = _numbers(va_path, vtext, 1, _INT32_MAX, "weight").reshape(m, k) if m else np.zeros((0, k), np.int64)
    return row, col.astype(np.int32), np.ascontiguousarray(w.astype(np.int32))


def read_csr_triplet(prefix, *, num_weights: int = 0, properties: GraphProperties | str | None = None,
                     resources: Resources | None = None) -> Graph:
    """Read MOSP's text CSR into a device graph (``properties="mosp_compatible"``: MOSP's batch
    semantics, parallel edges kept)."""
    if isinstance(properties, str):
        presets = {"mosp_compatible": GraphProperties.mosp_compatible, "default": GraphProperties}
        if properties not in presets:
            raise InvalidArgumentError(f"read_csr_triplet: unknown properties preset {properties!r}")
        properties = presets[properties]()
    row, col, w = read_csr_arrays(prefix, num_weights=num_weights)
    return Graph.from_csr(row, col, w, properties=properties, resources=resources)


def _write(path, text: str) -> None:
    p = Path(path)
    p.parent.mkdir(parents=True, exist_ok=True)
    with open(p, "w", encoding="ascii", newline="\n") as f:
        f.write(text)


def _join_rows(columns: list[np.ndarray], sep: str = " ") -> str:
    """One line per row of the integer columns."""
    if not columns or len(columns[0]) == 0:
        return ""
    parts = [c.astype(str) for c in columns]
    out = parts[0]
    for p in parts[1:]:
        out = np.char.add(np.char.add(out, sep), p)
    return "\n".join(out.tolist()) + "\n"


def write_csr_triplet(prefix, graph: Graph) -> None:
    """Write the graph's out-rows as MOSP's text CSR (writeCsrGraph's bytes: one offset, one
    neighbour, one line of K weights per line). Rows come out in the device graph's order: by
    neighbour id, parallel edges in the order they were added."""
    row, col, w = graph.to_csr()  # w is (K, m)
    rp, ci, va = _csr_files(prefix)
    _write(rp, _join_rows([np.asarray(row, np.int64)]))
    _write(ci, _join_rows([np.asarray(col, np.int64)]))
    _write(va, _join_rows([np.asarray(w[k], np.int64) for k in range(w.shape[0])]))


# ------------------------------------------------------------------------------------------------
# Batches
# ------------------------------------------------------------------------------------------------


def read_legacy_batch(insert_path, delete_path, *, num_weights: int = 1, num_vertices: int = -1,
                      mosp_lenient: bool = False) -> EdgeBatch:
    """Read MOSP's ``insert.txt`` / ``delete.txt`` (cpp/src/io/batch_io.cpp:27-139): ids in
    [0, num_vertices) (num_vertices -1: ids >= 0), weights in [1, 2^31 - 1], operations in file
    order."""
    k = int(num_weights)
    if k < 0:
        raise InvalidArgumentError(f"read_legacy_batch: num_weights must be >= 0, got {k}")
    id_max = int(num_vertices) - 1 if num_vertices >= 0 else _INT32_MAX
    ipath, dpath = os.fspath(insert_path), os.fspath(delete_path)
    expected = 2 + k
    ins = []
    for number, fields in _lines(_text(ipath)):
        if len(fields) > expected:
            if not mosp_lenient:
                raise FileFormatError(ipath, number, f"an insertion line holds exactly {expected} integers "
                                                     f"'u v w1 .. wK' (K = {k})")
            for tok in fields[expected:]:  # read and ignored, as MOSP's readChangeBatch does
                _int(ipath, number, tok, -(2**63), 2**63 - 1, "integer")
        if len(fields) < expected:
            raise FileFormatError(ipath, number, f"an insertion line holds exactly {expected} integers "
                                                 f"'u v w1 .. wK' (K = {k}), found {len(fields)}")
        u = _int(ipath, number, fields[0], 0, id_max, "vertex id")
        v = _int(ipath, number, fields[1], 0, id_max, "vertex id")
        ws = [_int(ipath, number, t, 1, _INT32_MAX, "weight") for t in fields[2:expected]]
        ins.append([u, v] + ws)
    dele = []
    for number, fields in _lines(_text(dpath)):
        if len(fields) > 2:
            if not mosp_lenient:
                raise FileFormatError(dpath, number, "a deletion line holds exactly 2 integers 'u v'")
            for tok in fields[2:]:
                _int(dpath, number, tok, -(2**63), 2**63 - 1, "integer")
        if mosp_lenient:
            values = [_int(dpath, number, t, -(2**63), 2**63 - 1, "vertex id") for t in fields[:2]]
            if len(values) == 1:
                continue  # MOSP skips a deletion line with one integer
            for x in values:
                if x < 0 or x > id_max:
                    raise FileFormatError(dpath, number, f"vertex id {x} is out of range [0, {id_max}]")
            dele.append(values)
        else:
            if len(fields) != 2:
                raise FileFormatError(dpath, number, "a deletion line holds exactly 2 integers 'u v'")
            dele.append([_int(dpath, number, t, 0, id_max, "vertex id") for t in fields])
    ia = np.array(ins, dtype=np.int64).reshape(-1, expected)
    da = np.array(dele, dtype=np.int64).reshape(-1, 2)
    insert = (ia[:, 0].copy(), ia[:, 1].copy(), np.ascontiguousarray(ia[:, 2:].astype(np.int32)))
    return EdgeBatch(insert=insert, delete=(da[:, 0].copy(), da[:, 1].copy()))


def write_legacy_batch(insert_path, delete_path, batch: EdgeBatch) -> None:
    """Write a host batch as MOSP's ``insert.txt`` / ``delete.txt`` (writeChangeBatch's bytes)."""
    s, d, w = (np.asarray(x) for x in (batch.insert_src, batch.insert_dst, batch.insert_weights))
    w = w.reshape(len(s), -1) if len(s) else np.zeros((0, 0), np.int64)
    _write(insert_path, _join_rows([s.astype(np.int64), d.astype(np.int64)] +
                                   [w[:, k].astype(np.int64) for k in range(w.shape[1])]))
    ds, dd = np.asarray(batch.delete_src), np.asarray(batch.delete_dst)
    _write(delete_path, _join_rows([ds.astype(np.int64), dd.astype(np.int64)]))


# ------------------------------------------------------------------------------------------------
# Result files
# ------------------------------------------------------------------------------------------------


def _vertex_values(path, num_vertices: int, parse) -> list:
    p = os.fspath(path)
    n = int(num_vertices)
    if n < 0:
        raise InvalidArgumentError(f"num_vertices must be >= 0, got {n}")
    values = [None] * n
    listed = 0
    for number, fields in _lines(_text(p)):
        if len(fields) != 2:
            raise FileFormatError(p, number, "a line holds exactly 2 values 'v x'")
        v = _int(p, number, fields[0], 0, n - 1, "vertex id")
        if values[v] is not None:
            raise FileFormatError(p, number, f"vertex {v} is listed twice")
        values[v] = parse(p, number, fields[1])
        listed += 1
    if listed != n:
        raise FileFormatError(p, 0, f"the file must list each of the {n} vertices exactly once ({listed} listed)")
    return values


def read_distances(path, num_vertices: int) -> np.ndarray:
    """A distance file (``v d`` or ``v INF`` per vertex) as int64 (INF reads as
    INFINITE_DISTANCE)."""
    def parse(p, number, tok):
        return INFINITE_DISTANCE if tok == "INF" else _int(p, number, tok, 0, 2**63 - 1, "distance")
    return np.array(_vertex_values(path, num_vertices, parse), dtype=np.int64).reshape(-1)


def read_parents(path, num_vertices: int) -> np.ndarray:
    """A tree file (``v p`` per vertex, -1 for none) as int32."""
    n = int(num_vertices)
    return np.array(_vertex_values(path, n, lambda p, number, tok: _int(p, number, tok, -1, n - 1, "parent")),
                    dtype=np.int32).reshape(-1)


def _inf_column(x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, np.int64)
    return np.where(x >= INFINITE_DISTANCE // 2, "INF", x.astype(str))


def write_distances(path, distances: Any) -> None:
    """``v d`` per vertex, ``INF`` at or above INFINITE_DISTANCE / 2 (writeDistances' bytes)."""
    d = np.asarray(distances, np.int64).reshape(-1)
    _write(path, _join_rows([np.arange(len(d)), _inf_column(d)]))


def write_parents(path, parents: Any) -> None:
    """``v p`` per vertex (writeParents' bytes)."""
    p = np.asarray(parents, np.int64).reshape(-1)
    _write(path, _join_rows([np.arange(len(p)), p]))


def write_path_costs(path, costs: Any) -> None:
    """``v c1 .. cK`` per vertex, ``INF`` at or above INFINITE_DISTANCE / 2 (the bytes of the
    ``mosp`` driver's mospCosts.txt)."""
    c = np.asarray(costs, np.int64)
    if c.ndim != 2:
        raise InvalidArgumentError("write_path_costs: costs must be an (n, K) array")
    _write(path, _join_rows([np.arange(c.shape[0])] + [_inf_column(c[:, k]) for k in range(c.shape[1])]))
