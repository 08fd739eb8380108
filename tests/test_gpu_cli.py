"""The command line (python -m dyng_amd) against the reference's committed files, byte for byte
(python/tests/test_cli.py:312-339, 689-721, 783-810): `sssp compute` writes the `mospPrep init`
trees, `sssp update` (from the original trees and from computed ones) the `mosp` driver's updated
trees, `mosp update --init` the combined files of every preference case, `mosp compute` the
combined files the oracle gives. Needs an MI355X: every test here is marked gpu."""

from __future__ import annotations

import os
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle
from oracle import mosp as omosp
from tests import fixture_io as fio

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def cli():
    import dyng_amd
    from dyng_amd.cli import main

    dyng_amd._lib.lib()
    assert dyng_amd.device_count() > 0, "no HIP device visible"
    return lambda *argv: main([str(a) for a in argv])


def _same_tree(a: Path, b: Path) -> None:
    fa = sorted(p.relative_to(a) for p in a.rglob("*") if p.is_file())
    fb = sorted(p.relative_to(b) for p in b.rglob("*") if p.is_file())
    assert fa == fb, (a, b)
    for f in fa:
        assert (a / f).read_bytes() == (b / f).read_bytes(), f


@pytest.mark.parametrize("name,inp,K", fio.sssp_cases(), ids=[c[0] for c in fio.sssp_cases()])
def test_sssp_compute_and_update_equal_mosp(cli, tmp_path, name, inp, K):
    _say(f"[cli sssp] {name}")
    expected = Path(fio.GOLDEN) / "mosp_sssp" / name
    graph = ["--graph", Path(inp) / "graphCsr", "--num-weights", K]
    assert cli("sssp", "compute", *graph, "--out", tmp_path / "c", "--quiet") == 0
    _same_tree(tmp_path / "c", expected / "init")
    for i, init in enumerate((["--init", expected / "init"], [])):
        out = tmp_path / f"u{i}"
        assert cli("sssp", "update", *graph, "--changes", inp, *init, "--out", out, "--quiet") == 0
        _same_tree(out, expected / "updated")


@pytest.mark.parametrize("name,inp,init,K,prefs", fio.combined_cases(), ids=[c[0] for c in fio.combined_cases()])
def test_mosp_update_equals_the_originals_combined_files(cli, tmp_path, capsys, name, inp, init, K, prefs):
    _say(f"[cli mosp] {name}")
    out = tmp_path / name
    status = cli("mosp", "update", "--graph", Path(inp) / "graphCsr", "--changes", inp, "--init", init,
                 "--num-objectives", K, "--preferences", ",".join(map(str, prefs)), "--out", out)
    text = capsys.readouterr().out
    assert status == 0
    assert "combined: L=" in text and "update " in text
    _same_tree(out / "combinedGraph", Path(fio.GOLDEN) / "mosp_combined" / name / "combined")
    assert sorted(p.name for p in out.iterdir()) == ["combinedGraph"] + [f"obj{i}" for i in range(K)]


@pytest.mark.parametrize("name,inp,K", fio.sssp_cases()[:8], ids=[c[0] for c in fio.sssp_cases()[:8]])
def test_mosp_compute_writes_the_trees_and_the_combined_files(cli, tmp_path, capsys, name, inp, K):
    prefs = [2] + [1] * (K - 1)
    out = tmp_path / name
    assert cli("mosp", "compute", "--graph", Path(inp) / "graphCsr", "--num-weights", K, "--out", out,
               "--preferences", ",".join(map(str, prefs))) == 0
    assert capsys.readouterr().out.startswith("compute ")
    for i in range(K):
        _same_tree(out / f"obj{i}", Path(fio.GOLDEN) / "mosp_sssp" / name / "init" / f"obj{i}")
    row, col, w = fio.read_csr(os.path.join(inp, "graphCsr"), K)
    g = oracle.Csr(row, col, w)
    trees = [oracle.dijkstra(g, 0, k)[1] for k in range(K)]
    d, p, _ = omosp.solve(trees, 0, prefs)
    from dyng_amd import io

    want = tmp_path / "want"
    io.write_distances(want / "distancesCsr.txt", d)
    io.write_parents(want / "SSSPTreeCsr.txt", p)
    io.write_path_costs(want / "mospCosts.txt", omosp.path_costs(g, p, 0, K))
    _same_tree(out / "combinedGraph", want)


def test_fewer_objectives_than_columns_and_write_graph(cli, tmp_path):
    """--num-objectives below the graph's columns: mospCosts.txt still covers every column;
    --write-graph writes the updated graph (sorted rows here) as the oracle's apply gives it."""
    name, inp, init, K, prefs = [c for c in fio.combined_cases() if c[0] == "r04_k2"][0]
    out = tmp_path / "o"
    assert cli("mosp", "update", "--graph", Path(inp) / "graphCsr", "--changes", inp, "--init", init,
               "--num-objectives", K, "--preferences", ",".join(map(str, prefs)), "--out", out, "--quiet",
               "--properties", "default", "--write-graph", tmp_path / "g_") == 0
    _same_tree(out / "combinedGraph", Path(fio.GOLDEN) / "mosp_combined" / name / "combined")
    cols = fio.num_weight_columns(os.path.join(inp, "graphCsr"))
    row, col, w = fio.read_csr(os.path.join(inp, "graphCsr"), cols)
    b = fio.read_batch(inp, cols)
    g1 = oracle.apply_batch(oracle.Csr(row, col, w), *b).g
    from dyng_amd import io

    r2, c2, w2 = io.read_csr_arrays(tmp_path / "g_")
    rows = np.repeat(np.arange(g1.n), np.diff(g1.row_ptr))
    order = np.lexsort((g1.col, rows))  # a sorted graph's rows, by neighbour id
    assert np.array_equal(r2, g1.row_ptr) and np.array_equal(c2, g1.col[order])
    assert np.array_equal(w2, g1.w[:, order].T)


def test_cli_errors(cli, tmp_path, capsys):
    name, inp, K = fio.sssp_cases()[0]
    assert cli("sssp", "update", "--graph", Path(inp) / "graphCsr", "--out", tmp_path) == 2  # no batch
    assert "a batch is required" in capsys.readouterr().err
    assert cli("sssp", "compute", "--graph", tmp_path / "absent_", "--out", tmp_path) == 1
    assert cli("sssp", "compute", "--graph", Path(inp) / "graphCsr", "--source", 10**6, "--out", tmp_path) == 1
    assert cli("mosp", "compute", "--graph", Path(inp) / "graphCsr", "--num-objectives", 99,
               "--out", tmp_path) == 1
