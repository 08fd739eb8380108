"""The ROCm plugin's native module called as dynG's typed layer calls its extension module
(python/dyng/sssp.py:346,385; python/bindings/sssp.cpp:77-135): a graph built with
GraphI32I32I32.from_edges, sssp_compute, batches of EdgeBatchI32I32 through sssp_update, and
sssp_from_arrays; trees, stats and rows against the oracle. Needs an MI355X: marked gpu.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    import dyng_amd
    import dyng_amd.rocm_plugin as plugin

    dyng_amd._lib.lib()  # loud failure if the HIP library is missing
    assert dyng_amd.device_count() > 0, "no HIP device visible"
    return plugin.native


def test_the_typed_layer_call_sequence(native):
    rng = np.random.default_rng(77)
    n, m = 3000, 30000
    src, dst = rng.integers(0, n, m).astype(np.int32), rng.integers(0, n, m).astype(np.int32)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    w = rng.integers(1, 100, len(src)).astype(np.int32)
    og = oracle.build_from_edges(n, src, dst, w.reshape(-1, 1))
    res = native.Resources.cuda(0, None, 0)
    props = native.GraphProperties()
    props.store_transposed = True
    g = native.GraphI32I32I32.from_edges(res, n, src, dst, w, 1, props)
    assert (g.num_vertices, g.num_edges, g.num_weights, g.space) == (n, og.m, 1, native.MemorySpace.device)
    opt = native.SsspOptions()
    r = native.sssp_compute(res, g, 0, opt)
    ot, _ = oracle.compute(og, 0)
    assert np.array_equal(r.distances(), ot.dist) and np.array_equal(r.parents(), ot.parent)
    assert r.source == 0 and r.graph_version == g.version
    for b in range(3):
        rows = np.repeat(np.arange(og.n), np.diff(og.row_ptr))
        pick = rng.choice(og.m, 200, replace=False)
        ins = (rng.integers(0, n, 200).astype(np.int32), rng.integers(0, n, 200).astype(np.int32),
               rng.integers(1, 100, 200).astype(np.int32))
        dels = (rows[pick].astype(np.int32), og.col[pick].astype(np.int32))
        og, a, c = oracle.update(og, ot, ins[0], ins[1], ins[2].reshape(-1, 1), dels[0], dels[1])
        batch = native.EdgeBatchI32I32(ins[0], ins[1], ins[2], dels[0], dels[1], 1, None, None, None)
        st = native.sssp_update(res, g, batch, r)
        assert np.array_equal(r.distances(), ot.dist) and np.array_equal(r.parents(), ot.parent), f"batch {b}"
        assert (st.invalidated, st.affected) == (c.invalidated, c.affected)
        assert st.engine_used == native.Engine.fused and st.packed_parents
        for f in ("inserted_edges", "updated_edges", "deleted_edges", "ignored_deletions"):
            assert getattr(st.batch, f) == a.summary[f], f
        assert r.generation == b + 1
    row, col, wk = g.to_csr(res)
    assert np.array_equal(row, og.row_ptr) and np.array_equal(col, og.col) and np.array_equal(wk, og.w)
    # an imported tree, and a clone
    r2 = native.sssp_from_arrays(res, g, 0, ot.dist, ot.parent, False, native.SsspOptions())
    assert np.array_equal(r2.parents(), ot.parent)
    r3 = r.clone(res)
    assert np.array_equal(r3.distances(), ot.dist)


def test_errors_of_the_typed_layer(native):
    from dyng_amd import errors

    res = native.Resources()
    props = native.GraphProperties()
    g = native.GraphI32I32I32.from_edges(res, 4, np.array([0, 1, 2], np.int32), np.array([1, 2, 3], np.int32),
                                         np.array([1, 1, 1], np.int32), 1, props)
    with pytest.raises(errors.InvalidArgumentError, match="out of range"):
        native.sssp_compute(res, g, 9, native.SsspOptions())
    r = native.sssp_compute(res, g, 0, native.SsspOptions())
    i32 = np.int32
    g.apply(res, native.EdgeBatchI32I32(np.array([3], i32), np.array([0], i32), np.array([2], i32),
                                        np.array([], i32), np.array([], i32), 1, None, None, None))
    with pytest.raises(errors.StaleResultError):  # the graph moved on without this result
        native.sssp_update(res, g, native.EdgeBatchI32I32(np.array([], i32), np.array([], i32), None,
                                                          np.array([], i32), np.array([], i32), 1, None, None,
                                                          None), r)
