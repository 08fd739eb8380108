"""CPU oracle of dynG's batch-dynamic SSSP path. TEST INFRASTRUCTURE ONLY.

A ctypes view of ``oracle/liboracle.so`` (``oracle/oracle.c``, a C restatement of the
reference's sequential engine, OpenMP engine, MOSP batch apply, transpose, Dijkstra and tree
check; file:line citations are in ``oracle.c``), plus the per-objective preparation of
``cpp/src/algorithms/sssp/sssp.cpp:643-689`` (largest weight, default delta) so a test can run
``compute`` / ``update`` exactly as ``dyng::sssp`` does on the host backends.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
this package; the product (``dyng_amd``) never does.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

INF = (2**63 - 1) // 4

_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)


class _Csr(C.Structure):
    _fields_ = [("n", C.c_int64), ("m", C.c_int64), ("K", C.c_int32), ("row_ptr", _i64p),
                ("col", _i32p), ("w", _i32p)]


class _Props(C.Structure):
    _fields_ = [(name, C.c_int32) for name in (
        "directed", "sorted_rows", "allow_parallel", "on_existing_insert", "on_missing_delete",
        "on_self_loop", "deletions_first", "allow_vertex_growth")]


_SUMMARY_FIELDS = ("inserted_edges", "updated_edges", "deleted_edges", "ignored_deletions",
                   "dropped_self_loops", "inserted_vertices", "num_vertices_after",
                   "ignored_insertions")


class _Summary(C.Structure):
    _fields_ = [(name, C.c_int64) for name in _SUMMARY_FIELDS]


class _Applied(C.Structure):
    _fields_ = [("g", _Csr), ("num_ins", C.c_int64), ("ins_src", _i32p), ("ins_dst", _i32p),
                ("weight_increased", _u8p), ("num_del", C.c_int64), ("del_src", _i32p),
                ("del_dst", _i32p), ("summary", _Summary)]


class _Run(C.Structure):
    _fields_ = [("n", C.c_int64), ("out_row", _i64p), ("out_col", _i32p), ("out_w", _i32p),
                ("in_row", _i64p), ("in_col", _i32p), ("in_w", _i32p),
                ("changed_from", _i32p), ("changed_to", _i32p), ("num_changed", C.c_int64),
                ("insert_heads", _i32p), ("num_insert_heads", C.c_int64),
                ("source", C.c_int32), ("threads", C.c_int32), ("delta", C.c_int64),
                ("max_weight", C.c_int64), ("dist", _i64p), ("parent", _i32p),
                ("invalidated", C.c_int64), ("iterations", C.c_int64), ("epochs", C.c_int64),
                ("pushes", C.c_int64), ("affected", C.c_int64), ("packed_parents", C.c_int32),
                ("t_hooks", C.c_double * 4)]


class OracleError(Exception):
    """An error the restated engines report (the reference throws invalid_argument_error)."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


_lib = None


def build() -> str:
    """Compile liboracle.so with oracle/Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_ws_new.restype = C.c_void_p
        L.orc_ws_free.argtypes = [C.c_void_p]
        L.orc_build_from_edges.argtypes = [C.c_int64, C.c_int32, C.c_int64, _i32p, _i32p, _i32p,
                                           C.POINTER(_Props), C.POINTER(_Csr), C.c_char_p, C.c_int]
        L.orc_apply_batch.argtypes = [C.POINTER(_Csr), C.POINTER(_Props), C.c_int64, _i32p, _i32p,
                                      _i32p, C.c_int64, _i32p, _i32p, C.c_int32,
                                      C.POINTER(_Applied), C.c_char_p, C.c_int]
        L.orc_transpose.argtypes = [C.POINTER(_Csr), C.POINTER(_Csr), C.c_int32]
        L.orc_csr_free.argtypes = [C.POINTER(_Csr)]
        L.orc_applied_free.argtypes = [C.POINTER(_Applied)]
        L.orc_default_delta.restype = C.c_int64
        L.orc_default_delta.argtypes = [C.c_int64, C.c_int64, C.c_int64]
        for name in ("orc_distances_fit", "orc_packs_parents", "orc_device_parent_bits"):
            getattr(L, name).restype = C.c_int32
            getattr(L, name).argtypes = [C.c_int64, C.c_int64]
        for name in ("orc_seq_compute", "orc_seq_update", "orc_omp_compute", "orc_omp_update"):
            getattr(L, name).argtypes = [C.POINTER(_Run), C.c_void_p, C.c_char_p, C.c_int]
        L.orc_dijkstra.argtypes = [C.c_int64, _i64p, _i32p, _i32p, C.c_int32, _i64p, _i32p]
        L.orc_check_tree.argtypes = [C.c_int64, _i64p, _i32p, _i32p, C.c_int32, _i64p, _i32p,
                                     C.c_int32, _i64p]
        L.orc_canonicalize.argtypes = [C.c_int64, _i64p, _i32p, _i32p, C.c_int32, _i64p, _i32p]
        _lib = L
    return _lib


def _p32(a: np.ndarray) -> C.POINTER(C.c_int32):
    return a.ctypes.data_as(_i32p)


def _p64(a: np.ndarray) -> C.POINTER(C.c_int64):
    return a.ctypes.data_as(_i64p)


def _arr32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.int64).astype(np.int32))


@dataclass
class Props:
    """graph_properties / batch_semantics switches (graph_properties.hpp:67-221)."""

    directed: bool = True
    sorted_rows: bool = True
    allow_parallel: bool = False
    on_existing_insert: int = 0  # 0 upsert, 1 error, 2 ignore
    on_missing_delete: int = 0  # 0 ignore, 1 error
    on_self_loop: int = 0  # 0 keep, 1 drop, 2 error
    deletions_first: bool = True
    allow_vertex_growth: bool = True

    @staticmethod
    def mosp_compatible() -> "Props":
        return Props(sorted_rows=False, allow_parallel=True)

    def _c(self) -> _Props:
        return _Props(int(self.directed), int(self.sorted_rows), int(self.allow_parallel),
                      self.on_existing_insert, self.on_missing_delete, self.on_self_loop,
                      int(self.deletions_first), int(self.allow_vertex_growth))


@dataclass
class Csr:
    """Compact rows; ``w`` is (K, m), objective-major."""

    row_ptr: np.ndarray
    col: np.ndarray
    w: np.ndarray

    @property
    def n(self) -> int:
        return len(self.row_ptr) - 1

    @property
    def m(self) -> int:
        return len(self.col)

    @property
    def K(self) -> int:
        return self.w.shape[0]

    def _c(self) -> _Csr:
        self.row_ptr = np.ascontiguousarray(self.row_ptr, dtype=np.int64)
        self.col = np.ascontiguousarray(self.col, dtype=np.int32)
        self.w = np.ascontiguousarray(self.w, dtype=np.int32)
        return _Csr(self.n, self.m, self.K, _p64(self.row_ptr), _p32(self.col),
                    self.w.ctypes.data_as(_i32p))

    @staticmethod
    def _from_c(c: _Csr) -> "Csr":
        n, m, K = c.n, c.m, c.K
        row = np.ctypeslib.as_array(c.row_ptr, shape=(n + 1,)).copy()
        col = np.ctypeslib.as_array(c.col, shape=(max(m, 1),))[:m].copy()
        w = np.ctypeslib.as_array(c.w, shape=(max(m * K, 1),))[: m * K].copy().reshape(K, m)
        return Csr(row, col, w)


@dataclass
class Applied:
    g: Csr
    ins_src: np.ndarray
    ins_dst: np.ndarray
    weight_increased: np.ndarray  # (num_ins, K)
    del_src: np.ndarray
    del_dst: np.ndarray
    summary: dict = field(default_factory=dict)

    def changes(self, k: int) -> tuple[np.ndarray, np.ndarray]:
        """build_changes (sssp.cpp:563-576): deletions, then weight-increased insertions."""
        inc = self.weight_increased[:, k].astype(bool) if len(self.ins_src) else np.zeros(0, bool)
        return (np.concatenate([self.del_src, self.ins_src[inc]]).astype(np.int32),
                np.concatenate([self.del_dst, self.ins_dst[inc]]).astype(np.int32))


def _err_buf():
    return C.create_string_buffer(512)


def build_from_edges(n: int, src, dst, weights, props: Props | None = None) -> Csr:
    props = props or Props()
    s, d = _arr32(src), _arr32(dst)
    w = np.ascontiguousarray(np.asarray(weights, dtype=np.int64).reshape(len(s), -1).astype(np.int32))
    K = w.shape[1] if w.ndim == 2 else 1
    out = _Csr()
    err = _err_buf()
    cp = props._c()
    rc = lib().orc_build_from_edges(n, K, len(s), _p32(s), _p32(d), w.ctypes.data_as(_i32p),
                                    C.byref(cp), C.byref(out), err, 512)
    if rc != 0:
        raise OracleError(rc, err.value.decode())
    try:
        return Csr._from_c(out)
    finally:
        lib().orc_csr_free(C.byref(out))


def apply_batch(g: Csr, ins_src, ins_dst, ins_w, del_src, del_dst,
                props: Props | None = None, threads: int = 1) -> Applied:
    props = props or Props()
    s, d = _arr32(ins_src), _arr32(ins_dst)
    w = np.ascontiguousarray(np.asarray(ins_w, dtype=np.int64).reshape(len(s), -1).astype(np.int32)) \
        if len(s) else np.zeros((0, g.K), np.int32)
    ds, dd = _arr32(del_src), _arr32(del_dst)
    out = _Applied()
    err = _err_buf()
    cg = g._c()
    cp = props._c()
    rc = lib().orc_apply_batch(C.byref(cg), C.byref(cp), len(s), _p32(s), _p32(d),
                               w.ctypes.data_as(_i32p), len(ds), _p32(ds), _p32(dd), threads,
                               C.byref(out), err, 512)
    if rc != 0:
        raise OracleError(rc, err.value.decode())
    try:
        ni, nd, K = out.num_ins, out.num_del, g.K
        return Applied(
            g=Csr._from_c(out.g),
            ins_src=np.ctypeslib.as_array(out.ins_src, shape=(ni + 1,))[:ni].copy(),
            ins_dst=np.ctypeslib.as_array(out.ins_dst, shape=(ni + 1,))[:ni].copy(),
            weight_increased=np.ctypeslib.as_array(out.weight_increased, shape=(ni * K + 1,))[
                : ni * K].copy().reshape(ni, K),
            del_src=np.ctypeslib.as_array(out.del_src, shape=(nd + 1,))[:nd].copy(),
            del_dst=np.ctypeslib.as_array(out.del_dst, shape=(nd + 1,))[:nd].copy(),
            summary={f: getattr(out.summary, f) for f in _SUMMARY_FIELDS},
        )
    finally:
        lib().orc_applied_free(C.byref(out))


def transpose(g: Csr) -> Csr:
    out = _Csr()
    cg = g._c()
    lib().orc_transpose(C.byref(cg), C.byref(out), 1)
    try:
        return Csr._from_c(out)
    finally:
        lib().orc_csr_free(C.byref(out))


def default_delta(m: int, n: int, weight_sum: int) -> int:
    return int(lib().orc_default_delta(m, n, weight_sum))


def distances_fit(n: int, max_weight: int) -> bool:
    return bool(lib().orc_distances_fit(n, max_weight))


def packs_parents(n: int, max_weight: int) -> bool:
    return bool(lib().orc_packs_parents(n, max_weight))


def device_parent_bits(n: int, max_weight: int) -> int:
    return int(lib().orc_device_parent_bits(n, max_weight))


def dijkstra(g: Csr, source: int, k: int = 0) -> tuple[np.ndarray, np.ndarray]:
    cg = g._c()
    w = np.ascontiguousarray(g.w[k])
    dist = np.empty(g.n, np.int64)
    par = np.empty(g.n, np.int32)
    lib().orc_dijkstra(g.n, _p64(g.row_ptr), _p32(g.col), _p32(w), source, _p64(dist), _p32(par))
    del cg
    return dist, par


def check_tree(g: Csr, source: int, dist, parent, k: int = 0,
               require_canonical: bool = False) -> dict:
    """check_sssp_tree (check_sssp.cpp:33-125)."""
    g._c()
    w = np.ascontiguousarray(g.w[k])
    d = np.ascontiguousarray(dist, dtype=np.int64)
    p = np.ascontiguousarray(parent, dtype=np.int32)
    counts = np.zeros(4, np.int64)
    lib().orc_check_tree(g.n, _p64(g.row_ptr), _p32(g.col), _p32(w), source, _p64(d), _p32(p),
                         int(require_canonical), _p64(counts))
    return {"distance_mismatches": int(counts[0]), "inconsistent_parents": int(counts[1]),
            "non_canonical_parents": int(counts[2]), "parent_mismatches": int(counts[3]),
            "ok": int(counts[0]) == 0 and int(counts[1]) == 0
            and (not require_canonical or int(counts[3]) == 0)}


def canonicalize(g: Csr, source: int, dist, parent, k: int = 0) -> np.ndarray:
    g._c()
    w = np.ascontiguousarray(g.w[k])
    d = np.ascontiguousarray(dist, dtype=np.int64)
    p = np.array(parent, dtype=np.int32)
    lib().orc_canonicalize(g.n, _p64(g.row_ptr), _p32(g.col), _p32(w), source, _p64(d), _p32(p))
    return p


@dataclass
class Counters:
    invalidated: int = 0
    iterations: int = 0
    epochs: int = 0
    pushes: int = 0
    affected: int = 0
    packed_parents: bool = True
    t_hooks: tuple = (0.0, 0.0, 0.0, 0.0)


class Workspace:
    """The engines' pooled workspace (sized once, reused; ADR 0015)."""

    def __init__(self):
        self.ptr = C.c_void_p(lib().orc_ws_new())

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.orc_ws_free(self.ptr)
            self.ptr = None


def run_engine(engine: str, update: bool, out: Csr, inn: Csr | None, k: int, source: int,
               delta: int, max_weight: int, dist: np.ndarray, parent: np.ndarray,
               changed_from=None, changed_to=None, insert_heads=None, threads: int = 1,
               ws: Workspace | None = None) -> Counters:
    """One engine run in place on (dist, parent); engine is 'seq' or 'omp'."""
    out._c()
    ow = np.ascontiguousarray(out.w[k]) if out.m else np.zeros(1, np.int32)
    r = _Run()
    r.n = out.n
    r.out_row, r.out_col, r.out_w = _p64(out.row_ptr), _p32(out.col), _p32(ow)
    keep = [ow]
    if inn is not None:
        inn._c()
        iw = np.ascontiguousarray(inn.w[k]) if inn.m else np.zeros(1, np.int32)
        keep.append(iw)
        r.in_row, r.in_col, r.in_w = _p64(inn.row_ptr), _p32(inn.col), _p32(iw)
    cf = _arr32(changed_from if changed_from is not None else [])
    ct = _arr32(changed_to if changed_to is not None else [])
    ih = _arr32(insert_heads if insert_heads is not None else [])
    keep += [cf, ct, ih]
    r.changed_from, r.changed_to, r.num_changed = _p32(cf), _p32(ct), len(cf)
    r.insert_heads, r.num_insert_heads = _p32(ih), len(ih)
    r.source, r.threads, r.delta, r.max_weight = source, threads, delta, max_weight
    assert dist.dtype == np.int64 and parent.dtype == np.int32
    assert dist.flags.c_contiguous and parent.flags.c_contiguous
    r.dist, r.parent = _p64(dist), _p32(parent)
    fn = {("seq", False): "orc_seq_compute", ("seq", True): "orc_seq_update",
          ("omp", False): "orc_omp_compute", ("omp", True): "orc_omp_update"}[(engine, update)]
    err = _err_buf()
    wsp = (ws or Workspace())
    rc = getattr(lib(), fn)(C.byref(r), wsp.ptr, err, 512)
    if rc != 0:
        raise OracleError(rc, err.value.decode())
    return Counters(r.invalidated, r.iterations, r.epochs, r.pushes, r.affected,
                    bool(r.packed_parents), tuple(r.t_hooks))


# --------------------------------------------------------------------------------------------
# The sssp flow of the host backends (sssp.cpp:643-764, 1208-1249), on oracle graphs.
# --------------------------------------------------------------------------------------------


@dataclass
class Tree:
    source: int
    dist: np.ndarray
    parent: np.ndarray
    objective: int = 0
    delta: int = 0  # options::delta (0 = automatic)


def column_summary(g: Csr, k: int) -> tuple[int, int]:
    """summarize_column (sssp.cpp:285-318): (max weight >= 1, weight sum)."""
    col = g.w[k].astype(np.int64) if g.m else np.zeros(0, np.int64)
    return (max(1, int(col.max())) if len(col) else 1), int(col.sum())


def compute(g: Csr, source: int, k: int = 0, engine: str = "seq", delta: int = 0,
            threads: int = 1, ws: Workspace | None = None) -> tuple[Tree, Counters]:
    n = g.n
    if not (0 <= source < n):
        raise OracleError(1, f"sssp::compute: source {source} is out of range [0, {n})")
    max_w, wsum = column_summary(g, k)
    if not distances_fit(n, max_w):
        raise OracleError(1, "sssp::compute: distances do not fit in 62 bits")
    d = delta if delta > 0 else default_delta(max(g.m, 1), n, wsum)
    dist = np.full(n, INF, np.int64)
    par = np.full(n, -1, np.int32)
    inn = transpose(g) if engine == "seq" else None
    c = run_engine(engine, False, g, inn, k, source, d, max_w, dist, par, threads=threads, ws=ws)
    return Tree(source, dist, par, k, delta), c


def update(g: Csr, tree: Tree, ins_src, ins_dst, ins_w, del_src, del_dst,
           props: Props | None = None, engine: str = "seq", threads: int = 1,
           ws: Workspace | None = None) -> tuple[Csr, Applied, Counters]:
    """sssp::update on the host backends: prepare on G_t, commit, resume, the engine's hooks."""
    props = props or Props()
    k = tree.objective
    ins_w = np.asarray(ins_w, dtype=np.int64).reshape(len(np.atleast_1d(ins_src)), -1) \
        if len(np.atleast_1d(ins_src)) else np.zeros((0, g.K), np.int64)
    if len(ins_w) and (ins_w[:, k] < 1).any():
        raise OracleError(1, "sssp::update: an insertion has a weight below 1")
    max_w, wsum = column_summary(g, k)
    largest = int(ins_w[:, k].max()) if len(ins_w) else 1
    max_weight = max(largest, max_w)
    delta = tree.delta if tree.delta > 0 else default_delta(max(g.m, 1), g.n, wsum)
    applied = apply_batch(g, ins_src, ins_dst, ins_w, del_src, del_dst, props)
    g2 = applied.g
    if not distances_fit(g2.n, max_weight):
        raise OracleError(1, "sssp::update: distances do not fit in 62 bits")
    if g2.n > len(tree.dist):
        grow = g2.n - len(tree.dist)
        tree.dist = np.concatenate([tree.dist, np.full(grow, INF, np.int64)])
        tree.parent = np.concatenate([tree.parent, np.full(grow, -1, np.int32)])
    cf, ct = applied.changes(k)
    inn = transpose(g2)
    c = run_engine(engine, True, g2, inn, k, tree.source, delta, max_weight, tree.dist,
                   tree.parent, cf, ct, applied.ins_dst, threads=threads, ws=ws)
    return g2, applied, c


def update_many(g: Csr, trees: list, ins_src, ins_dst, ins_w, del_src, del_dst,
                props: Props | None = None, engine: str = "seq", threads: int = 1,
                workspaces: list | None = None) -> tuple[Csr, Applied, list]:
    """dyng::update_each on the host backends (cpp/include/dyng/update.hpp:311-364): every tree's
    preparation on G_t, the batch applied once, then each tree repaired on G_{t+1} (one
    transpose shared by all). Returns (G_{t+1}, the apply, one Counters per tree)."""
    props = props or Props()
    m_in = len(np.atleast_1d(ins_src))
    ins_w = np.asarray(ins_w, dtype=np.int64).reshape(m_in, -1) if m_in else np.zeros((0, g.K), np.int64)
    prepared = []
    for tree in trees:
        k = tree.objective
        if len(ins_w) and (ins_w[:, k] < 1).any():
            raise OracleError(1, "sssp::update: an insertion has a weight below 1")
        max_w, wsum = column_summary(g, k)
        largest = int(ins_w[:, k].max()) if len(ins_w) else 1
        delta = tree.delta if tree.delta > 0 else default_delta(max(g.m, 1), g.n, wsum)
        prepared.append((max(largest, max_w), delta))
    applied = apply_batch(g, ins_src, ins_dst, ins_w, del_src, del_dst, props)
    g2 = applied.g
    inn = transpose(g2)
    out = []
    for i, (tree, (max_weight, delta)) in enumerate(zip(trees, prepared)):
        k = tree.objective
        if not distances_fit(g2.n, max_weight):
            raise OracleError(1, "sssp::update: distances do not fit in 62 bits")
        if g2.n > len(tree.dist):
            grow = g2.n - len(tree.dist)
            tree.dist = np.concatenate([tree.dist, np.full(grow, INF, np.int64)])
            tree.parent = np.concatenate([tree.parent, np.full(grow, -1, np.int32)])
        cf, ct = applied.changes(k)
        ws = workspaces[i] if workspaces is not None else None
        out.append(run_engine(engine, True, g2, inn, k, tree.source, delta, max_weight, tree.dist,
                              tree.parent, cf, ct, applied.ins_dst, threads=threads, ws=ws))
    return g2, applied, out
