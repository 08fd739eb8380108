"""dyng_amd.io (MOSP's text CSR, batch and result files) against the reference's committed files
(cpp/tests/data/mosp_sssp, mosp_combined, copied as data under tests/golden/): the readers give the
fixture's arrays, the writers give back the fixture's bytes; the strict and the lenient (MOSP)
batch rules of cpp/src/io/batch_io.cpp:27-139; malformed files name the file and line. No GPU."""

from __future__ import annotations

import os

import numpy as np
import pytest

from dyng_amd import EdgeBatch, io
from dyng_amd.cli import build_parser, main

from . import fixture_io as fio


def _bytes(path) -> bytes:
    with open(path, "rb") as f:
        return f.read()


@pytest.mark.parametrize("name,inp,K", fio.sssp_cases(), ids=[c[0] for c in fio.sssp_cases()])
def test_readers_and_writers_match_the_fixtures(tmp_path, name, inp, K):
    row, col, w = io.read_csr_arrays(os.path.join(inp, "graphCsr"))
    r2, c2, w2 = fio.read_csr(os.path.join(inp, "graphCsr"), K)
    assert np.array_equal(row, r2) and np.array_equal(col, c2) and np.array_equal(w, w2.T)
    b = io.read_legacy_batch(os.path.join(inp, "insert.txt"), os.path.join(inp, "delete.txt"), num_weights=K,
                             num_vertices=len(row) - 1, mosp_lenient=True)
    e = fio.read_batch(inp, K)
    assert np.array_equal(b.insert_src, e[0]) and np.array_equal(b.insert_dst, e[1])
    assert np.array_equal(b.insert_weights, e[2]) and np.array_equal(b.delete_src, e[3])
    assert np.array_equal(b.delete_dst, e[4])
    n = len(row) - 1
    for stage, suffix in (("init", "Original"), ("updated", "Updated")):
        for k in range(K):
            d = os.path.join(fio.GOLDEN, "mosp_sssp", name, stage, f"obj{k}")
            dist = io.read_distances(os.path.join(d, f"distances{suffix}.txt"), n)
            par = io.read_parents(os.path.join(d, f"SSSPTree{suffix}.txt"), n)
            io.write_distances(tmp_path / "d.txt", dist)
            io.write_parents(tmp_path / "p.txt", par)
            assert _bytes(tmp_path / "d.txt") == _bytes(os.path.join(d, f"distances{suffix}.txt"))
            assert _bytes(tmp_path / "p.txt") == _bytes(os.path.join(d, f"SSSPTree{suffix}.txt"))


@pytest.mark.parametrize("case", fio.combined_cases(), ids=[c[0] for c in fio.combined_cases()])
def test_combined_files_round_trip(tmp_path, case):
    name = case[0]
    d = os.path.join(fio.GOLDEN, "mosp_combined", name, "combined")
    n = len(fio.read_text(os.path.join(d, "SSSPTreeCsr.txt")).splitlines())
    io.write_distances(tmp_path / "a.txt", io.read_distances(os.path.join(d, "distancesCsr.txt"), n))
    io.write_parents(tmp_path / "b.txt", io.read_parents(os.path.join(d, "SSSPTreeCsr.txt"), n))
    io.write_path_costs(tmp_path / "c.txt", fio.read_costs(os.path.join(d, "mospCosts.txt"), n))
    assert _bytes(tmp_path / "a.txt") == _bytes(os.path.join(d, "distancesCsr.txt"))
    assert _bytes(tmp_path / "b.txt") == _bytes(os.path.join(d, "SSSPTreeCsr.txt"))
    assert _bytes(tmp_path / "c.txt") == _bytes(os.path.join(d, "mospCosts.txt"))


def _files(tmp_path, ins: str, dele: str):
    (tmp_path / "insert.txt").write_text(ins)
    (tmp_path / "delete.txt").write_text(dele)
    return tmp_path / "insert.txt", tmp_path / "delete.txt"


def test_lenient_and_strict_batch_rules(tmp_path):
    i, d = _files(tmp_path, "0 1 5 7 99\n\n2 3 4 4\n", "0 1 8\n5\n2 3\n")
    b = io.read_legacy_batch(i, d, num_weights=2, num_vertices=10, mosp_lenient=True)
    assert b.insert_src.tolist() == [0, 2] and b.insert_weights.tolist() == [[5, 7], [4, 4]]
    assert b.delete_src.tolist() == [0, 2] and b.delete_dst.tolist() == [1, 3]  # "5" skipped, "8" ignored
    with pytest.raises(io.FileFormatError) as e:
        io.read_legacy_batch(i, d, num_weights=2, num_vertices=10)
    assert e.value.line == 1 and "insert.txt" in str(e.value)
    i, d = _files(tmp_path, "0 1 5 7\n", "0 1\n5\n")
    with pytest.raises(io.FileFormatError) as e:
        io.read_legacy_batch(i, d, num_weights=2, num_vertices=10)
    assert e.value.line == 2
    i, d = _files(tmp_path, "0 1 0 7\n", "")
    with pytest.raises(io.FileFormatError, match="weight"):
        io.read_legacy_batch(i, d, num_weights=2, mosp_lenient=True)
    i, d = _files(tmp_path, "0 12 1 7\n", "")
    with pytest.raises(io.FileFormatError, match="out of range"):
        io.read_legacy_batch(i, d, num_weights=2, num_vertices=10, mosp_lenient=True)
    i, d = _files(tmp_path, "0 1 1\n", "0 1x\n")
    with pytest.raises(io.FileFormatError, match="not an integer"):
        io.read_legacy_batch(i, d, num_weights=1, mosp_lenient=True)


def test_batch_round_trip(tmp_path):
    b = EdgeBatch(insert=([3, 1], [5, 5], [[10, 2, 12], [12, 1, 14]]), delete=([1, 4], [3, 1]))
    io.write_legacy_batch(tmp_path / "insert.txt", tmp_path / "delete.txt", b)
    assert (tmp_path / "insert.txt").read_text() == "3 5 10 2 12\n1 5 12 1 14\n"
    assert (tmp_path / "delete.txt").read_text() == "1 3\n4 1\n"
    r = io.read_legacy_batch(tmp_path / "insert.txt", tmp_path / "delete.txt", num_weights=3)
    assert r.insert_weights.tolist() == [[10, 2, 12], [12, 1, 14]] and r.delete_dst.tolist() == [3, 1]


def test_malformed_csr_and_result_files(tmp_path):
    p = tmp_path / "g"
    (tmp_path / "gRowPtr.txt").write_text("0\n2\n1\n")
    (tmp_path / "gColInd.txt").write_text("1\n0\n")
    (tmp_path / "gValues.txt").write_text("1 2\n3 4\n")
    with pytest.raises(io.FileFormatError, match="must not decrease"):
        io.read_csr_arrays(p)
    (tmp_path / "gRowPtr.txt").write_text("0\n1\n2\n")
    (tmp_path / "gValues.txt").write_text("1 2\n3\n")
    with pytest.raises(io.FileFormatError) as e:
        io.read_csr_arrays(p)
    assert e.value.line == 2
    (tmp_path / "gValues.txt").write_text("1 2\n3 0\n")
    with pytest.raises(io.FileFormatError, match="weight 0"):
        io.read_csr_arrays(p)
    (tmp_path / "gValues.txt").write_text("1 2\n3 4\n")
    row, col, w = io.read_csr_arrays(p)
    assert row.tolist() == [0, 1, 2] and col.tolist() == [1, 0] and w.tolist() == [[1, 2], [3, 4]]
    (tmp_path / "d.txt").write_text("0 0\n1 INF\n1 3\n")
    with pytest.raises(io.FileFormatError, match="listed twice"):
        io.read_distances(tmp_path / "d.txt", 2)
    (tmp_path / "d.txt").write_text("0 0\n")
    with pytest.raises(io.FileFormatError, match="exactly once"):
        io.read_distances(tmp_path / "d.txt", 2)
    with pytest.raises(io.FileFormatError, match="cannot read"):
        io.read_parents(tmp_path / "absent.txt", 2)
    io.write_path_costs(tmp_path / "c.txt", np.array([[0, 0], [io.INFINITE_DISTANCE, 5]]))
    assert (tmp_path / "c.txt").read_text() == "0 0 0\n1 INF 5\n"


def test_cli_usage_errors():
    assert main(["sssp"]) == 2  # no verb
    assert main(["mosp", "update", "--graph", "x"]) == 2  # no --out
    args = build_parser().parse_args(["mosp", "update", "--graph", "g", "--changes", "c", "--preferences", "4,1,4",
                                      "--num-objectives", "3", "--out", "o"])
    assert args.preferences == "4,1,4" and args.num_objectives == 3 and args.func.__name__ == "_run_mosp_update"
