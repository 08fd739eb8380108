"""Every benchmark configuration at its own size through the HIP path, against the oracle.

  C1  ER, 65,536 vertices, 1,048,576 edges, w ~ U[1, 100], source 0; 10 batches of 500 new-edge
      insertions + 500 deletions; every batch against the sequential engine (the reference's
      sequential backend restated).
  C2  RMAT-22 (4,194,304 vertices, 67,108,864 edges), w ~ U[1, 255], source = largest out-degree;
      3 mixed batches of 100,000 operations (40 % insertions, 40 % deletions, 20 % reweights);
      against the OpenMP engine.
  C3  4096 x 4096 grid (16,777,216 vertices, 4-neighbour arcs both ways, 0.1 % shortcuts),
      w ~ U[1, 1000], source at the corner; 2 batches of 10,000 operations drawn against the
      current tree (80 % of the deletions on tree edges); against the OpenMP engine.
  C4  24 independent RMAT instances (n = 2^10 ... 2^19, average degree 16, w ~ U[1, 100]) as one
      union graph: one apply and one persistent launch per step (the last steps replayed from a
      hipGraph); 3 steps of each instance's own 1 % batch; every instance against the OpenMP engine
      run on the instance alone.
  C5  RMAT-23 (8,388,608 vertices, 134,217,728 edges), w ~ U[1, 255]: a slice of 20 stream batches
      of log-uniform size in [10, 10^6] and uniform delete fraction (including one of more than
      10^5 operations), the plain schedule and the learned scheduler in lockstep, each against the
      OpenMP engine after every batch.

The inputs come from the project's generators (splitmix64 draws, the seeds of the benchmark).
Distances, parents, `invalidated` and `affected` must be equal after every batch. Needs an
MI355X: every test here is marked gpu.
"""

from __future__ import annotations

import os
import sys
import time

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dg():
    import dyng_amd

    dyng_amd._lib.lib()  # loud failure if the HIP library is missing
    assert dyng_amd.device_count() > 0, "no HIP device visible"
    return dyng_amd


def _csr(n, src, dst, w):
    """The oracle graph of sorted distinct edges (the generators' output order)."""
    row = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(src, minlength=n), out=row[1:])
    return oracle.Csr(row, dst.astype(np.int32), np.asarray(w, np.int32).reshape(1, -1))


def _dev(dg, x):
    D = dg.DeviceArray
    return dg.EdgeBatch(insert=(D(x[0]), D(x[1]), D(x[2])), delete=(D(x[3]), D(x[4])))


def _check(r, ot, st, c, what):
    d, p = r.to_numpy()
    assert np.array_equal(d, ot.dist), f"{what}: distances"
    assert np.array_equal(p, ot.parent), f"{what}: parents"
    if st is not None:
        assert (st.invalidated, st.affected) == (c.invalidated, c.affected), f"{what}: invalidated/affected"


@pytest.mark.timeout(600)
def test_c1_er64k_every_batch_against_the_sequential_engine(dg):
    from dyng_amd import generators as G

    n, m = 65_536, 1_048_576
    src, dst, w = G.er(n, m, 100, G.GRAPH_SEED ^ 0)
    batches = G.mixed_batches(src, dst, n, 10, 500, 500, 0, 100, G.BATCH_SEED ^ 0)
    _say(f"[C1] ER n={n} m={m}, 10 x 1000 ops")
    og = _csr(n, src, dst, w)
    hg = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    ot, _ = oracle.compute(og, 0)
    r = dg.sssp.compute(hg, 0)
    _check(r, ot, None, None, "compute")
    for b, x in enumerate(batches):
        og, _, c = oracle.update(og, ot, x[0], x[1], x[2].reshape(-1, 1), x[3], x[4])
        st = dg.sssp.update(hg, _dev(dg, x), r)
        _check(r, ot, st, c, f"batch {b}")
        assert st.packed_parents


@pytest.mark.timeout(600)
def test_c2_rmat22_mixed100k_against_the_openmp_engine(dg):
    from dyng_amd import generators as G

    scale, m = 22, 67_108_864
    n = 1 << scale
    t0 = time.perf_counter()
    src, dst, w = G.rmat(scale, m, 255, G.GRAPH_SEED ^ 1)
    batches = G.mixed_batches(src, dst, n, 3, 40_000, 40_000, 20_000, 255, G.BATCH_SEED ^ 1, rmat_scale=scale)
    source = G.max_out_degree_vertex(src, n)
    _say(f"[C2] RMAT-22 n={n} m={m} source={source}, inputs in {time.perf_counter() - t0:.1f}s")
    hg = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    r = dg.sssp.compute(hg, source)
    og = _csr(n, src, dst, w)
    del src, dst, w
    ws = oracle.Workspace()
    ot, _ = oracle.compute(og, source, 0, "omp", threads=THREADS, ws=ws)
    _check(r, ot, None, None, "compute")
    for b, x in enumerate(batches):
        og, _, c = oracle.update(og, ot, x[0], x[1], x[2].reshape(-1, 1), x[3], x[4], engine="omp", threads=THREADS,
                                 ws=ws)
        st = dg.sssp.update(hg, _dev(dg, x), r)
        _check(r, ot, st, c, f"batch {b}")
        _say(f"[C2] batch {b}: invalidated {st.invalidated}, affected {st.affected}, bit-exact")


@pytest.mark.timeout(600)
def test_c3_grid4096_tree_batches_against_the_openmp_engine(dg):
    from dyng_amd import generators as G

    side = 4096
    n = side * side
    t0 = time.perf_counter()
    src, dst, w = G.grid(side, 0.001, 1000, G.GRAPH_SEED ^ 2)
    _say(f"[C3] grid {side}x{side} n={n} m={len(src)}, inputs in {time.perf_counter() - t0:.1f}s")
    hg = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    r = dg.sssp.compute(hg, 0)
    og = _csr(n, src, dst, w)
    del src, dst, w
    ws = oracle.Workspace()
    ot, _ = oracle.compute(og, 0, 0, "omp", threads=THREADS, ws=ws)
    _check(r, ot, None, None, "compute")
    for b in range(2):
        x = G.tree_batch(side, ot.parent, 5_000, 4_000, 1_000, 1000, G.BATCH_SEED ^ 2, b)
        og, _, c = oracle.update(og, ot, x[0], x[1], x[2].reshape(-1, 1), x[3], x[4], engine="omp", threads=THREADS,
                                 ws=ws)
        st = dg.sssp.update(hg, _dev(dg, x), r)
        _check(r, ot, st, c, f"batch {b}")
        assert st.invalidated > 1000
        _say(f"[C3] batch {b}: invalidated {st.invalidated}, affected {st.affected}, bit-exact")


@pytest.mark.timeout(600)
def test_c4_batched_instances_one_launch_against_each_instance_alone(dg):
    from concurrent.futures import ThreadPoolExecutor

    from dyng_amd import batched
    from dyng_amd import generators as G

    cfg = 3
    scales = list(range(10, 20)) * 2 + [19, 18, 17, 16]

    def make(i):
        k = scales[i]
        n_i = 1 << k
        src, dst, w = G.rmat(k, 16 * n_i, 100, (G.GRAPH_SEED ^ cfg) + 7919 * (i + 1))
        half = max(1, len(src) // 200)
        drawn = G.mixed_batches(src, dst, n_i, 3, half, half, 0, 100, (G.BATCH_SEED ^ cfg) + 7919 * (i + 1),
                                rmat_scale=k)
        return (n_i, src, dst, w), G.max_out_degree_vertex(src, n_i), drawn

    with ThreadPoolExecutor(max_workers=THREADS) as pool:
        made = list(pool.map(make, range(len(scales))))
    graphs = [x[0] for x in made]
    sources = [x[1] for x in made]
    drawn = [x[2] for x in made]
    _say(f"[C4] {len(scales)} instances, n={sum(g[0] for g in graphs)} m={sum(len(g[1]) for g in graphs)}")
    inst = batched.Instances(graphs)
    r = inst.compute(sources)
    oracles = []
    for i, (n_i, s, d, w) in enumerate(graphs):
        og = _csr(n_i, s, d, w)
        ot, _ = oracle.compute(og, sources[i], 0, "omp", threads=THREADS)
        oracles.append([og, ot])
    arrays = r.to_numpy()
    for i, (og, ot) in enumerate(oracles):
        d, p = inst.tree(r, i, arrays)
        assert np.array_equal(d, ot.dist) and np.array_equal(p, ot.parent), f"instance {i} (2^{scales[i]}) compute"
    for b in range(3):
        if b == 1:
            r.set_tuning(dg.sssp.Tuning(launch="graph"))  # steps 1 and 2: the hipGraph replay
        st = inst.update([drawn[i][b] for i in range(len(scales))], r)
        inv = aff = 0
        for i in range(len(scales)):
            x = drawn[i][b]
            oracles[i][0], _, c = oracle.update(oracles[i][0], oracles[i][1], x[0], x[1], x[2].reshape(-1, 1), x[3],
                                                x[4], engine="omp", threads=THREADS)
            inv += c.invalidated
            aff += c.affected
        arrays = r.to_numpy()
        for i, (og, ot) in enumerate(oracles):
            d, p = inst.tree(r, i, arrays)
            assert np.array_equal(d, ot.dist) and np.array_equal(p, ot.parent), f"instance {i} (2^{scales[i]}) step {b}"
        assert (st.invalidated, st.affected) == (inv, aff), f"step {b}"
        _say(f"[C4] step {b}: one launch, {len(scales)} instances bit-exact")


@pytest.mark.timeout(900)
def test_c5_rmat23_stream_slice_plain_and_learned_against_the_openmp_engine(dg):
    from dyng_amd import generators as G
    from dyng_amd import scheduler, sssp

    scale, m, cfg = 23, 134_217_728, 4
    n = 1 << scale
    t0 = time.perf_counter()
    src, dst, w = G.rmat(scale, m, 255, G.GRAPH_SEED ^ cfg)
    source = G.max_out_degree_vertex(src, n)
    stream = G.stream(src, dst, n, 20, 10, 1_000_000, 255, G.BATCH_SEED ^ cfg, scale)
    train = G.stream(src, dst, n, 24, 10, 1_000_000, 255, (G.BATCH_SEED ^ cfg) ^ 0x7EA1, scale)
    sizes = [len(x[0]) + len(x[3]) for x in stream]
    assert max(sizes) > 100_000, sizes
    _say(f"[C5] RMAT-23 n={n} m={m}, 20 stream batches of {min(sizes)}..{max(sizes)} ops, inputs in "
         f"{time.perf_counter() - t0:.1f}s")
    # the scheduler, trained on traces of held-out batches (exploration on a graph of its own)
    gt = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    rt = sssp.compute(gt, source)
    trace = scheduler.explore(gt, rt, (_dev(dg, x) for x in train))
    del rt, gt
    sched = scheduler.Scheduler()
    sched.fit(trace, steps=500)
    gp = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    gl = dg.Graph.from_edges(src, dst, w, num_vertices=n)
    rp = sssp.compute(gp, source)
    rl = sssp.compute(gl, source)
    og = _csr(n, src, dst, w)
    del src, dst, w
    ws = oracle.Workspace()
    ot, _ = oracle.compute(og, source, 0, "omp", threads=THREADS, ws=ws)
    _check(rp, ot, None, None, "compute (plain)")
    _check(rl, ot, None, None, "compute (learned)")
    for b, x in enumerate(stream):
        og, _, c = oracle.update(og, ot, x[0], x[1], x[2].reshape(-1, 1), x[3], x[4], engine="omp", threads=THREADS,
                                 ws=ws)
        eb = _dev(dg, x)
        sp = sssp.update(gp, eb, rp)
        sl = sched.update(gl, eb, rl)
        _check(rp, ot, sp, c, f"plain, batch {b} ({sizes[b]} ops)")
        _check(rl, ot, sl, c, f"learned, batch {b} ({sizes[b]} ops)")
    _say(f"[C5] 20 batches bit-exact, learned picks {np.bincount(sched.choices, minlength=len(scheduler.CHOICES))}")
