"""The batch semantics' error modes on the device apply (graph_properties.hpp:71-83:
existing_insert::error, missing_delete::error), against the oracle's restated host apply
(apply_host.cpp:80-252): the same batches fail, the same operation is named, nothing changes
(strong guarantee), and error-free batches apply exactly as the oracle does. Needs an MI355X:
every test here is marked gpu.
"""

from __future__ import annotations

import re
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dg():
    import dyng_amd

    dyng_amd._lib.lib()  # loud failure if the HIP library is missing
    assert dyng_amd.device_count() > 0, "no HIP device visible"
    return dyng_amd


def _pair(dg, n, src, dst, w, directed, existing, missing, parallel):
    codes_ins = {"upsert": 0, "error": 1, "ignore": 2}
    codes_del = {"ignore": 0, "error": 1}
    op = oracle.Props(directed=directed, allow_parallel=parallel, on_existing_insert=codes_ins[existing],
                      on_missing_delete=codes_del[missing])
    sem = dg.BatchSemantics(on_existing_insert=existing, on_missing_delete=missing)
    gp = dg.GraphProperties(directed=directed, parallel_edges="allow" if parallel else "forbid", semantics=sem)
    og = oracle.build_from_edges(n, src, dst, np.asarray(w).reshape(-1, 1), op)
    hg = dg.Graph.from_edges(src, dst, w, num_vertices=n, properties=gp)
    return op, og, hg


_PAIR = re.compile(r"\((-?\d+), (-?\d+)\)")


@pytest.mark.parametrize("directed", [True, False])
@pytest.mark.parametrize("parallel", [False, True])
def test_error_modes_match_the_oracle(dg, directed, parallel):
    rng = np.random.default_rng(40 + 2 * int(directed) + int(parallel))
    n, m = 400, 3000
    src, dst, w = rng.integers(0, n, m), rng.integers(0, n, m), rng.integers(1, 30, m)
    keep = src != dst
    src, dst, w = src[keep], dst[keep], w[keep]
    failed = passed = 0
    for mode in [("error", "ignore"), ("upsert", "error"), ("error", "error")]:
        op, og, hg = _pair(dg, n, src, dst, w, directed, *mode, parallel)
        ot, _ = oracle.compute(og, 0)
        r = dg.sssp.compute(hg, 0)
        for b in range(12):
            rows = np.repeat(np.arange(og.n), np.diff(og.row_ptr))
            k = int(rng.integers(1, 40))
            pick = rng.choice(og.m, k, replace=False)
            dels = [rows[pick], og.col[pick].copy()]
            ni = int(rng.integers(1, 40))
            ins = [rng.integers(0, n, ni), rng.integers(0, n, ni), rng.integers(1, 30, ni)]
            fault = b % 3
            if fault == 1:  # one deletion of a pair that is not there (maybe past the vertex count)
                dels[0] = np.append(dels[0], rng.integers(0, n + 3))
                dels[1] = np.append(dels[1], rng.integers(0, n + 3))
            elif fault == 2:  # one insertion of an existing pair
                e = int(rng.integers(0, og.m))
                ins[0] = np.append(ins[0], rows[e])
                ins[1] = np.append(ins[1], og.col[e])
                ins[2] = np.append(ins[2], 7)
            ins[0], ins[1] = np.where(ins[0] == ins[1], (ins[1] + 1) % n, ins[0]), ins[1]
            batch = dg.EdgeBatch(insert=tuple(ins), delete=tuple(dels))
            try:
                og2, a, c = oracle.update(og, ot, ins[0], ins[1], ins[2].reshape(-1, 1), dels[0], dels[1], props=op)
                oracle_error = None
            except oracle.OracleError as e:
                oracle_error = str(e)
            version = hg.version
            before = hg.to_csr()
            if oracle_error is not None:
                with pytest.raises(dg.InvalidArgumentError) as info:
                    dg.sssp.update(hg, batch, r)
                got = _PAIR.findall(str(info.value))
                want = _PAIR.findall(oracle_error)
                assert got and want and got[-1] == want[-1], (str(info.value), oracle_error)
                assert hg.version == version
                after = hg.to_csr()
                assert all(np.array_equal(x, y) for x, y in zip(before, after)), "the graph changed"
                d, p = r.to_numpy()  # the result is still current
                assert np.array_equal(d, ot.dist) and np.array_equal(p, ot.parent)
                failed += 1
            else:
                st = dg.sssp.update(hg, batch, r)
                og = og2
                d, p = r.to_numpy()
                assert np.array_equal(d, ot.dist) and np.array_equal(p, ot.parent)
                assert (st.invalidated, st.affected) == (c.invalidated, c.affected)
                for f in ("inserted_edges", "updated_edges", "deleted_edges", "ignored_deletions",
                          "ignored_insertions"):
                    assert getattr(st.batch, f) == a.summary[f], f
                passed += 1
    _say(f"[semantics] directed={directed} parallel={parallel}: {failed} batches refused, {passed} applied")
    assert failed >= 6 and passed >= 6
