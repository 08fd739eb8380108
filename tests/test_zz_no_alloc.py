"""dyng_hip_sssp_update_noalloc (no counterpart in the reference: the update for callers that
cannot allocate) against dyng_hip_sssp_update with a workspace, through the C ABI: calls in a
row, the second on a smaller graph and the third on one of the second's size, each bit-exact
with the workspace entry point (and with the oracle); a graph too large for the static scratch
is refused before anything runs. Collected last (its name), so that it runs after every other
GPU test as well as alone. Needs an MI355X: marked gpu.
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


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a[:2], b[:2])) and \
        (a[2].invalidated, a[2].affected) == (b[2].invalidated, b[2].affected)


def test_noalloc_update_matches_the_workspace_entry_point_twice(dg):
    from dyng_amd import _lib

    L = _lib.lib()
    # the second call is smaller: the scratch keeps the first call's contents (its stamps, far
    # marks and flags now lie where the first call's words were); the third has the second's size
    for n, m, ops, seed in [(60000, 600000, 8000, 21), (4000, 30000, 500, 22), (4000, 30000, 500, 24)]:
        case = seam.make_case(dg, n, m, ops, seed)
        with seam.workspace(dg) as ws:
            plain = seam.run_update(dg, case, L.dyng_hip_sssp_update, ws)
        noalloc = seam.run_update(dg, case, L.dyng_hip_sssp_update_noalloc)
        assert _same(plain, noalloc), f"n={n}: the no-allocation update differs from the workspace one"
        assert np.array_equal(noalloc[0], case.want.dist) and np.array_equal(noalloc[1], case.want.parent)
        assert (noalloc[2].invalidated, noalloc[2].affected) == (case.counts.invalidated, case.counts.affected)
        assert noalloc[2].tree_token == 0 and noalloc[2].words_kept == 0


def test_noalloc_update_refuses_a_graph_past_its_scratch(dg):
    from dyng_amd import _lib

    n = _lib.NOALLOC_SCRATCH_BYTES // 88 + 1000  # 88 B per vertex alone is past the array
    case = seam.make_case(dg, n, 2 * n, 100, 23)
    dist, par = dg.DeviceArray(case.old_dist), dg.DeviceArray(case.old_parent)
    run = seam.fill_run(dg, case, dist, par)
    counters = _lib.Counters()
    with pytest.raises(dg.InvalidArgumentError, match="bytes of scratch"):
        _lib.check(_lib.lib().dyng_hip_sssp_update_noalloc(C.byref(run), None, C.byref(counters)))
    assert np.array_equal(dist.to_numpy(), case.old_dist) and np.array_equal(par.to_numpy(), case.old_parent)
