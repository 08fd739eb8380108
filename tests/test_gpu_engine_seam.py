"""The engine seam called directly, as a `backend::hip` branch of `sssp_problem::resume`
(sssp.cpp:691-823) would call it: a graph changed by `dyng_hip_graph_apply` alone, a
`dyng_hip_sssp_run` filled by hand (the counterpart of `sssp_run`, problem.hpp:380-400) and
`dyng_hip_sssp_update` / `dyng_hip_sssp_compute` with a workspace of the caller's. Trees,
`invalidated` and `affected` against the oracle's update. Needs an MI355X: marked gpu.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests import seam

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dg():
    import dyng_amd

    dyng_amd._lib.lib()  # loud failure if the HIP library is missing
    assert dyng_amd.device_count() > 0, "no HIP device visible"
    return dyng_amd


@pytest.mark.parametrize("check_cycles", [0, 1])
@pytest.mark.parametrize("n,m,ops,seed", [(3000, 24000, 400, 11), (20000, 200000, 4000, 12)])
def test_update_through_the_seam_matches_the_oracle(dg, n, m, ops, seed, check_cycles):
    from dyng_amd import _lib

    case = seam.make_case(dg, n, m, ops, seed)
    with seam.workspace(dg) as ws:
        d, p, c = seam.run_update(dg, case, _lib.lib().dyng_hip_sssp_update, ws, check_cycles)
    assert np.array_equal(d, case.want.dist), "distances"
    assert np.array_equal(p, case.want.parent), "parents"
    assert (c.invalidated, c.affected) == (case.counts.invalidated, case.counts.affected)
    assert c.packed_parents == 1 and c.engine_used == 1  # DYNG_HIP_ENGINE_FUSED


def test_compute_through_the_seam_matches_the_oracle(dg):
    from dyng_amd import _lib

    case = seam.make_case(dg, 5000, 40000, 600, 13)
    n = len(case.old_dist)
    dist = dg.DeviceArray(np.zeros(n, np.int64))
    par = dg.DeviceArray(np.zeros(n, np.int32))
    run = seam.fill_run(dg, case, dist, par)
    with seam.workspace(dg) as ws:
        run.workspace = ws
        counters = _lib.Counters()
        _lib.check(_lib.lib().dyng_hip_sssp_compute(C.byref(run), None, C.byref(counters)))
    # the canonical tree of the graph after the batch is the updated tree (update == compute)
    assert np.array_equal(dist.to_numpy(), case.want.dist)
    assert np.array_equal(par.to_numpy(), case.want.parent)


def test_seam_refuses_a_run_without_a_workspace(dg):
    from dyng_amd import _lib

    case = seam.make_case(dg, 500, 3000, 40, 14)
    dist, par = dg.DeviceArray(case.old_dist), dg.DeviceArray(case.old_parent)
    run = seam.fill_run(dg, case, dist, par, workspace=None)
    counters = _lib.Counters()
    with pytest.raises(dg.InvalidArgumentError, match="workspace"):
        _lib.check(_lib.lib().dyng_hip_sssp_update(C.byref(run), None, C.byref(counters)))
    assert np.array_equal(dist.to_numpy(), case.old_dist)
