#!/usr/bin/env python3
"""bench.py -- edges/s and p50 ms/batch of dyng_amd.sssp.update on one MI355X.

Workload (the headline configuration): RMAT scale 22 (4,194,304 vertices, 67,108,864 distinct
directed non-loop edges, a=0.57 b=0.19 c=0.19 d=0.05), weights U[1,255], source = the vertex with
the largest out-degree; mixed batches of 100,000 operations (40% insertions of new edges with
RMAT endpoints, 40% deletions of existing edges, 20% reweights of existing edges to U[1,255]).
Graph seed 12345 ^ 1, batch seed 777 ^ 1 (splitmix64 counter-based draws).

A step is one sssp.update call (device batch apply + repair) on a batch already resident in HBM.
`value` = batch operations / p50 wall seconds per update. The JSON line also carries the
persistent kernel's roofline (algorithmic bytes / HIP-event duration against 8 TB/s), the
OpenMP CPU baseline (the oracle's restatement of the reference engine, timed on this host on a
bounded sample) and the bit-exactness check of the GPU trees against that CPU engine.
"""

from __future__ import annotations

import argparse
import dataclasses
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "edges/s & p50 ms/batch, RMAT-22 100K mixed batches; learned vs plain GPU"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (kind, size, edges, wmax, config index, batches, (ins, del, rew))
    "rmat22_mixed100k": ("rmat", 22, 67_108_864, 255, 1, 20, (40_000, 40_000, 20_000)),
    "er64k_mixed1k": ("er", 65_536, 1_048_576, 100, 0, 10, (500, 500, 0)),
    # grid side 4096 (16,777,216 vertices, 4-neighbour arcs both ways + 0.1 % shortcuts); each
    # batch is drawn against the current tree: (insertions, tree-edge deletions, random grid-arc
    # deletions), so 80 % of the deletions hit tree edges
    "grid4096_tree10k": ("grid", 4096, None, 1000, 2, 10, (5_000, 4_000, 1_000)),
    # RMAT-23 (8,388,608 vertices, 134,217,728 edges) and a stream of 10,000 batches of
    # log-uniform size in [10, 10^6] with a uniform delete fraction: the learned scheduler against
    # the plain one
    "rmat23_learned": ("stream", 23, 134_217_728, 255, 4, 10_000, (10, 1_000_000, 0)),
    # 256 independent RMAT graphs, n = 2^floor(U[10, 20)) each, average degree 16, weights U[1, 100],
    # one source each (its largest out-degree vertex); every step applies each instance's own batch
    # of 1 % of its edges (half insertions of new RMAT edges, half deletions) and repairs all 256
    # trees in one persistent launch replayed from a hipGraph, with autotuned schedule knobs
    "batched256_rmat": ("batched", 256, None, 100, 3, 6, (0, 0, 0)),
    # MOSP on RMAT-22 with three weight columns (column 0 as rmat22_mixed100k draws it, columns 1-2
    # U[1, 255] from their own seed), preferences (4, 1, 4): each step is one mosp.update (the
    # batch applied once, three trees repaired, the kept combined graph brought up to date and its
    # tree repaired, the path costs recomputed)
    "mosp_rmat22_k3": ("mosp", 22, 67_108_864, 255, 5, 20, (40_000, 40_000, 20_000)),
    # hyper_sssp: 200,000 hyperedges of 2-6 members drawn uniformly over 100,000 vertices, weights
    # U[1, 100], the source the vertex in most hyperedges; each step is one hyper_sssp.update with
    # 500 hyperedge deletions, 500 insertions, 500 incidence removals and 500 additions
    "hyper_uniform": ("hyper", 100_000, 200_000, 100, None, 20, (500, 500, 500)),
}
MOSP_OBJECTIVES = 3
MOSP_PREFERENCES = [4, 1, 4]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(workload: str):
    """HBM bytes per persistent-kernel launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    passes of this workload (scaled by a copy of known size), or None if none was taken. A
    workload name ending in "_compute" names the passes over its solves from the source."""
    path = os.path.join(ROOT, "profiles", f"sssp_fused_kernel_traffic_{workload}.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch = "):
                    return int(line.split("=")[1])
    except OSError:
        pass
    return None


def fused_kernel_bytes(c, n: int, num_changes: int, num_ins: int) -> int:
    """Algorithmic HBM bytes of one persistent-kernel launch (update), from its work counters:
    packing the old tree, 24 B/vertex, unless the workspace kept its words; 20 B per
    pointer-jumping visit, 20 B per invalidated vertex, 12 B per change-list entry, 8 B per
    insertion head, 20 B per pull candidate plus 16 B per in-edge it scans, 20 B per expanded
    vertex + 16 B per scanned out-edge + 16 B per improving atomicMin; the unpack visits the
    vertices it stamped (24 B each: list entry, word, old pair; the invalidated ones and, at
    least, the changed ones) and writes 12 B per changed (distance, parent) pair."""
    per_vertex = 0 if c.words_kept else 24
    return (per_vertex * n + 20 * c.jump_items + 20 * c.invalidated + 12 * num_changes + 8 * num_ins
            + 20 * c.pull_vertices + 16 * c.pull_edges
            + 20 * c.relax_vertices + 16 * c.relax_edges + 16 * c.relax_improved
            + 24 * (c.invalidated + c.affected) + 12 * c.affected)


def compute_bytes(c, n: int) -> int:
    """Algorithmic HBM bytes of a solve from the source: every word set (8 B/vertex), the
    near-far propagation (20 P + 16 E_s + 16 I), every vertex unpacked (word read, distance and
    parent written: 20 B/vertex)."""
    return 8 * n + relax_bytes(c) + 20 * n


def relax_bytes(c) -> int:
    """B_relax = 20 P + 16 E_s + 16 I (the near-far propagation)."""
    return 20 * c.relax_vertices + 16 * c.relax_edges + 16 * c.relax_improved


def make_inputs(name: str, batches_needed: int):
    """The graph, the source and `batch(b, parents)`: batch b of the workload (mixed batches are
    drawn up front; tree batches against the parents of the tree they will update)."""
    from dyng_amd import generators as G

    kind, size, m, wmax, cfg, nb, (ni, nd, nr) = WORKLOADS[name]
    nb = max(nb, batches_needed)
    t0 = time.perf_counter()
    if kind == "grid":
        n = size * size
        src, dst, w = G.grid(size, 0.001, wmax, G.GRAPH_SEED ^ cfg)
        m = len(src)

        def batch(b, parents):
            return G.tree_batch(size, parents, ni, nd, nr, wmax, G.BATCH_SEED ^ cfg, b)

        batch.dynamic = True
    else:
        if kind == "rmat":
            n = 1 << size
            src, dst, w = G.rmat(size, m, wmax, G.GRAPH_SEED ^ cfg)
            drawn = G.mixed_batches(src, dst, n, nb, ni, nd, nr, wmax, G.BATCH_SEED ^ cfg, rmat_scale=size)
        else:
            n = size
            src, dst, w = G.er(n, m, wmax, G.GRAPH_SEED ^ cfg)
            drawn = G.mixed_batches(src, dst, n, nb, ni, nd, nr, wmax, G.BATCH_SEED ^ cfg)

        def batch(b, parents):
            return drawn[b]

        batch.dynamic = False
    source = G.max_out_degree_vertex(src, n) if kind == "rmat" else 0
    log(f"[bench] inputs {name}: n={n} m={m} batches={nb}x{ni + nd + nr} in {time.perf_counter() - t0:.1f}s")
    return n, src, dst, w, batch, source, ni + nd + nr


# The CPUs this process may run on, read at start: once an OpenMP region has run with
# OMP_PROC_BIND set (the generators' library runs one), the main thread is bound to one core and
# its own affinity shrinks to that core.
STARTUP_CPUS = sorted(os.sched_getaffinity(0))


def host_cpus() -> dict:
    """The host cores this process may run on: CPU model, logical CPUs, physical cores and the
    sockets they sit on (from /proc/cpuinfo and sysfs topology)."""
    cpus = STARTUP_CPUS
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    cores, sockets = {}, {}
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            with open(base + "physical_package_id") as f:
                pkg = int(f.read())
            with open(base + "core_id") as f:
                core = int(f.read())
        except OSError:
            pkg, core = 0, c
        cores.setdefault((pkg, core), c)
        sockets.setdefault(pkg, set()).add((pkg, core))
    one_socket = max(len(v) for v in sockets.values()) if sockets else len(cpus)
    return {"model": model, "logical_cpus": len(cpus), "physical_cores": len(cores), "sockets": len(sockets),
            "cores_one_socket": one_socket}


def cpu_baseline(n, src, dst, w, batch, source, sample: int, thread_counts):
    """The oracle's OpenMP engine (openmp.cpp restated) on `sample` batches, once per thread
    count: repair time per batch (identify_affected + seed + loop + finalize) and the full host
    update (MOSP apply + transpose + repair). Returns (summary per thread count, trees after each
    sampled batch of the first run, the sampled batches)."""
    import oracle

    row = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(src, minlength=n), out=row[1:])
    summaries, trees, drawn = {}, [], []
    for run, threads in enumerate(thread_counts):
        g = oracle.Csr(row, dst.astype(np.int32), w.astype(np.int32).reshape(1, -1))
        ws = oracle.Workspace()
        t0 = time.perf_counter()
        tree, _ = oracle.compute(g, source, 0, "omp", threads=threads, ws=ws)
        t_compute = time.perf_counter() - t0
        repair, full = [], []
        for b in range(sample):
            if run == 0:
                drawn.append(batch(b, tree.parent))
            ins_s, ins_d, ins_w, del_s, del_d = drawn[b]
            t1 = time.perf_counter()
            g, _, c = oracle.update(g, tree, ins_s, ins_d, ins_w.reshape(-1, 1), del_s, del_d, engine="omp",
                                    threads=threads, ws=ws)
            full.append(time.perf_counter() - t1)
            repair.append(sum(c.t_hooks))
            if run == 0:
                trees.append((tree.dist.copy(), tree.parent.copy(), c.invalidated, c.affected))
            log(f"[bench] cpu batch {b} ({threads} threads): repair {repair[-1] * 1e3:.1f} ms, full update "
                f"{full[-1] * 1e3:.0f} ms")
        summaries[threads] = {"compute_s": t_compute, "repair_ms": [x * 1e3 for x in repair],
                              "update_ms": [x * 1e3 for x in full]}
        del g, ws, tree
    return summaries, trees, drawn


def learned_vs_plain(n, src, dst, w, source, train, evaluate, hindsight=(), log_prefix="learned"):
    """Train the scheduler on traces of the `train` batches (held-out seeds) on one device graph:
    every schedule's engine time on clones of the tree for each batch (scheduler.collect). Then run the `evaluate` batches through two more graphs in
    lockstep, one with the plain fixed schedule and one with the scheduler, comparing the two trees
    on the device after every batch. Update times are the wall time of each sssp.update call (the
    scheduler's feature kernel and BF16 inference included). Then, over the `hindsight` batches,
    every schedule is timed on clones of the tree: the best possible per-batch pick bounds what
    any scheduler could gain with these schedules."""
    import dyng_amd
    from dyng_amd import scheduler, sssp

    D = dyng_amd.DeviceArray

    def dev(x):
        return dyng_amd.EdgeBatch(insert=(D(x[0]), D(x[1]), D(x[2])), delete=(D(x[3]), D(x[4])))

    t0 = time.perf_counter()
    gt = dyng_amd.Graph.from_edges(src, dst, w, num_vertices=n)
    rt = sssp.compute(gt, source)
    # every schedule timed on clones of the tree for each training batch (complete targets: a
    # single random schedule per batch left the network's predictions too noisy to leave the plain
    # schedule safely)
    trace = scheduler.collect(gt, rt, (dev(x) for x in train))
    del rt, gt
    sched = scheduler.Scheduler()
    loss = sched.fit(trace, steps=3000)
    tx, tt = trace.arrays()
    seen = ~np.isnan(tt)
    by_choice = {c: round(float(np.mean(2.0 ** tt[seen[:, i], i])), 4) if seen[:, i].any() else None
                 for i, c in enumerate(scheduler.CHOICES)}
    log(f"[bench] {log_prefix}: {len(train)} training traces, fit loss {loss:.4f}, mean engine ms by schedule "
        f"{by_choice}, in {time.perf_counter() - t0:.1f}s")
    gp = dyng_amd.Graph.from_edges(src, dst, w, num_vertices=n)
    gl = dyng_amd.Graph.from_edges(src, dst, w, num_vertices=n)
    rp = sssp.compute(gp, source)
    rl = sssp.compute(gl, source)
    plain_ms, learned_ms, diffs = [], [], 0
    ops = 0
    for i, x in enumerate(evaluate):
        if i and i % 1000 == 0:
            log(f"[bench] {log_prefix}: {i} batches, plain {sum(plain_ms):.1f} ms, learned {sum(learned_ms):.1f} ms")
        eb = dev(x)
        ops += len(x[0]) + len(x[3])
        dyng_amd.synchronize()
        t1 = time.perf_counter()
        sssp.update(gp, eb, rp)
        plain_ms.append((time.perf_counter() - t1) * 1e3)
        t1 = time.perf_counter()
        sched.update(gl, eb, rl)
        learned_ms.append((time.perf_counter() - t1) * 1e3)
        diffs += rp.differences(rl)
        del eb
    # the ceiling: every schedule timed on clones of the plain tree over `hindsight` more batches of
    # the stream; the best per batch, picked with hindsight, against the plain schedule and against
    # the network's picks (engine times: the apply is the same under every schedule)
    held, hist, net_picks = scheduler.Trace(), scheduler.History(), []
    for x in hindsight:
        eb = dev(x)
        feats = scheduler.features(scheduler.batch_stats(gp, eb, rp), hist)
        net_picks.append(scheduler.pick(sched.network.predict(feats[None, :])[0]))
        scheduler.collect(gp, rp, [eb], held)
        c = rp.counters
        hist = scheduler.History(c.invalidated, c.affected)
        del eb
    ceiling = None
    if held.t:
        t = 2.0 ** np.asarray(held.t, np.float64)
        ceiling = {"batches": len(t), "plain_engine_ms": round(float(t[:, 0].sum()), 3),
                   "best_engine_ms": round(float(t.min(1).sum()), 3),
                   "learned_engine_ms": round(float(t[np.arange(len(t)), net_picks].sum()), 3),
                   "best_vs_plain": round(float(t[:, 0].sum() / t.min(1).sum()), 4),
                   "best_by_schedule": np.bincount(t.argmin(1), minlength=len(scheduler.CHOICES)).tolist(),
                   "timed": "every schedule on clones of the tree per batch (scheduler.collect)"}
        log(f"[bench] {log_prefix}: hindsight over {len(t)} batches: plain {ceiling['plain_engine_ms']:.1f} ms, "
            f"best {ceiling['best_engine_ms']:.1f} ms ({ceiling['best_vs_plain']:.3f}x), network's picks "
            f"{ceiling['learned_engine_ms']:.1f} ms")
    picks = np.bincount(sched.choices, minlength=len(scheduler.CHOICES)).tolist()
    out = {"batches": len(plain_ms), "ops": ops, "plain_ms_total": round(sum(plain_ms), 3),
           "learned_ms_total": round(sum(learned_ms), 3),
           "speedup": round(sum(plain_ms) / max(sum(learned_ms), 1e-9), 4),
           "plain_p50_ms": round(statistics.median(plain_ms), 4), "learned_p50_ms": round(statistics.median(learned_ms), 4),
           "picks": dict(zip(scheduler.CHOICES, picks)), "train_traces": len(train),
           "train": "every schedule timed on clones of the tree for each training batch (scheduler.collect)",
           "train_engine_ms_mean_by_schedule": by_choice,
           "bit_exact": diffs == 0, "compared": "plain and learned trees on the device after every batch",
           "hindsight": ceiling}
    log(f"[bench] {log_prefix}: plain {out['plain_ms_total']:.1f} ms, learned {out['learned_ms_total']:.1f} ms, "
        f"speedup {out['speedup']:.3f}x, picks {out['picks']}, bit-exact {out['bit_exact']}")
    del rp, rl, gp, gl
    return out


def stream_inputs(workload: str, count: int, train_batches: int):
    """The learned-scheduler workload: its graph, source, training batches (held-out seeds) and a
    generator of `count` evaluated stream batches."""
    from dyng_amd import generators as G

    kind, scale, m, wmax, cfg, nb, (lo, hi, _) = WORKLOADS[workload]
    n = 1 << scale
    t0 = time.perf_counter()
    src, dst, w = G.rmat(scale, m, wmax, G.GRAPH_SEED ^ cfg)
    source = G.max_out_degree_vertex(src, n)
    train = G.stream(src, dst, n, train_batches, lo, hi, wmax, (G.BATCH_SEED ^ cfg) ^ 0x7EA1, scale)

    def evaluate():
        # drawn in chunks of 500 batches (a whole stream of large batches would not fit in host
        # memory at once); chunk c uses seed (batch seed ^ config) + c
        for c in range(0, count, 500):
            yield from G.stream(src, dst, n, min(500, count - c), lo, hi, wmax, (G.BATCH_SEED ^ cfg) + c // 500, scale)

    hindsight = G.stream(src, dst, n, 100, lo, hi, wmax, (G.BATCH_SEED ^ cfg) ^ 0x51D3, scale)
    log(f"[bench] inputs {workload}: n={n} m={m} train {len(train)} batches in {time.perf_counter() - t0:.1f}s; "
        f"evaluating {count} batches")
    return n, m, src, dst, w, source, train, evaluate, hindsight, (lo, hi)


def stream_bench(args) -> int:
    """The learned-scheduler configuration: RMAT-23 and a stream of uneven batches."""
    count = args.stream_batches or WORKLOADS[args.workload][5]
    n, m, src, dst, w, source, train, evaluate, hindsight, (lo, hi) = stream_inputs(args.workload, count,
                                                                                   args.train_batches)
    res = learned_vs_plain(n, src, dst, w, source, train, evaluate(), hindsight)
    out = {"metric": METRIC, "value": round(res["ops"] / (res["learned_ms_total"] / 1e3), 1), "unit": "edges/s",
           "n_gpus": 1, "steps": count, "warmup": 0,
           "ms_per_step": round(res["learned_ms_total"] / count, 4), "p50_ms": res["learned_p50_ms"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64", "data": "synthetic",
           "config": {"workload": args.workload, "vertices": n, "edges": m, "batch_ops": f"log-uniform [{lo}, {hi}]",
                      "source": int(source)},
           "learned_vs_plain": res, "roofline": None, "cpu_baseline": None}
    print(json.dumps(out), flush=True)
    return 0 if res["bit_exact"] else 3


def batched_bench(args) -> int:
    """The batched multi-instance configuration: one device graph holding 256 RMAT instances."""
    import dyng_amd
    from dyng_amd import batched, sssp
    from dyng_amd import generators as G

    kind, count, _, wmax, cfg, nb, _ = WORKLOADS[args.workload]
    K, W = args.steps, args.warmup
    nb = max(nb, K + W)
    t0 = time.perf_counter()
    rng = np.random.default_rng(G.GRAPH_SEED ^ cfg)
    scales = np.floor(rng.uniform(10.0, 20.0, count)).astype(int)

    def make(i):
        k = int(scales[i])
        n_i = 1 << k
        src, dst, w = G.rmat(k, 16 * n_i, wmax, (G.GRAPH_SEED ^ cfg) + 7919 * (i + 1))
        half = max(1, len(src) // 200)
        drawn_i = G.mixed_batches(src, dst, n_i, nb, half, half, 0, wmax, (G.BATCH_SEED ^ cfg) + 7919 * (i + 1),
                                  rmat_scale=k)
        return (n_i, src, dst, w), G.max_out_degree_vertex(src, n_i), drawn_i

    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(max_workers=min(16, len(STARTUP_CPUS))) as pool:
        made = list(pool.map(make, range(count)))  # the generators release the GIL
    graphs = [x[0] for x in made]
    sources = [x[1] for x in made]
    drawn = [x[2] for x in made]
    del made
    m_total = sum(len(g[1]) for g in graphs)
    n_total = sum(g[0] for g in graphs)
    log(f"[bench] inputs {args.workload}: {count} instances, n={n_total} m={m_total}, {nb} batches each, "
        f"in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    inst = batched.Instances(graphs)
    tree = inst.compute(sources)
    D = dyng_amd.DeviceArray

    def union_batch(b):
        parts = [[], [], [], [], []]
        for i in range(count):
            x = drawn[i][b]
            off = int(inst.offsets[i])
            for f, a in enumerate(x):
                parts[f].append(a if f == 2 else np.asarray(a, np.int64) + off)
        cat = [np.concatenate(a) for a in parts]
        return dyng_amd.EdgeBatch(insert=(D(cat[0].astype(np.int32)), D(cat[1].astype(np.int32)), D(cat[2])),
                                  delete=(D(cat[3].astype(np.int32)), D(cat[4].astype(np.int32))))

    dev = [union_batch(b) for b in range(W + K)]
    ops = [sum(len(drawn[i][b][0]) + len(drawn[i][b][3]) for i in range(count)) for b in range(W + K)]
    dyng_amd.synchronize()
    log(f"[bench] device union graph + compute of {count} trees in {time.perf_counter() - t0:.1f}s")
    # warm-up batches: the autotuner times its candidates on the first one (on clones of the tree)
    report = None
    for b in range(W):
        if b == 0:
            report = batched.autotune(inst.graph, dev[0], tree)
            tree.set_tuning(dataclasses.replace(report.tuning, launch="graph"))
            log(f"[bench] autotune: light_degree {report.tuning.light_degree}, chunk {report.tuning.chunk_edges}, "
                f"delta {report.delta} ({report.engine_ms:.3f} ms; {len(report.table)} candidates)")
        else:
            sssp.update(inst.graph, dev[b], tree)
    wall, stats = [], []
    for b in range(W, W + K):
        dyng_amd.synchronize()
        t1 = time.perf_counter()
        st = sssp.update(inst.graph, dev[b], tree)
        wall.append((time.perf_counter() - t1) * 1e3)
        stats.append(st)
    p50 = statistics.median(wall)
    opsK = ops[W:]
    value = statistics.median(o / (t / 1e3) for o, t in zip(opsK, wall))
    log(f"[bench] gpu: p50 {p50:.3f} ms/step ({statistics.median(opsK)} ops), {value:.3e} edges/s")
    kbytes = [fused_kernel_bytes(s.device, n_total, s.batch.deleted_edges + s.batch.ignored_deletions + o, o)
              for s, o in zip(stats, opsK)]
    kms = [s.device.kernel_ms for s in stats]
    achieved = (sum(kbytes) / len(kbytes)) / (statistics.mean(kms) / 1e3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(args.workload),
                "kernel": "sssp_fused_kernel", "kernel_ms_mean": round(statistics.mean(kms), 4),
                "algorithmic_bytes_mean": int(sum(kbytes) / len(kbytes))}
    # ---- parity and the CPU baseline on a sample of instances (the smallest and a spread) ----
    cpu, exact = None, None
    if not args.no_cpu and args.cpu_sample > 0:
        import oracle

        threads = args.cpu_threads or len(STARTUP_CPUS)
        order = np.argsort(scales, kind="stable")
        sample = sorted(set(order[:4].tolist() + order[len(order) // 2::max(1, len(order) // 8)][:4].tolist()))
        d_all, p_all = tree.to_numpy()
        ok, rep_s, rep_ops = True, 0.0, 0
        for i in sample:
            n_i, src, dst, w = graphs[i]
            row = np.zeros(n_i + 1, np.int64)
            np.cumsum(np.bincount(src, minlength=n_i), out=row[1:])
            og = oracle.Csr(row, dst.astype(np.int32), w.astype(np.int32).reshape(1, -1))
            ws = oracle.Workspace()
            ot, _ = oracle.compute(og, int(sources[i]), 0, "omp", threads=threads, ws=ws)
            for b in range(W + K):
                x = drawn[i][b]
                og, _, c = oracle.update(og, ot, x[0], x[1], x[2].reshape(-1, 1), x[3], x[4], engine="omp",
                                         threads=threads, ws=ws)
                if b >= W:
                    rep_s += sum(c.t_hooks)
                    rep_ops += len(x[0]) + len(x[3])
            di, pi = inst.tree(tree, i, (d_all, p_all))
            same = bool(np.array_equal(di, ot.dist) and np.array_equal(pi, ot.parent))
            ok = ok and same
            log(f"[bench] parity instance {i} (n=2^{scales[i]}): {'bit-exact' if same else 'MISMATCH'} after "
                f"{W + K} batches")
        exact = {"instances_checked": len(sample), "batches_each": W + K, "bit_exact": ok,
                 "against": "OpenMP engine (oracle), each instance alone"}
        cpu = {"value": round(rep_ops / max(rep_s, 1e-9), 1), "unit": "edges/s", "cores": threads, "kind": "port",
               "sample": f"{len(sample)} instances x {K} batches (their own 1 % batches), repair time of the OpenMP "
                         f"engine per instance (identify_affected + seed + loop + finalize)"}
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "edges/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(sum(wall) / K, 4), "p50_ms": round(p50, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int64", "data": "synthetic",
        "config": {"workload": args.workload, "instances": count, "vertices": int(n_total), "edges": int(m_total),
                   "batch_ops_p50": int(statistics.median(opsK)), "launch": "hipGraph replay, one launch per step",
                   "autotuned": {"light_degree": report.tuning.light_degree if report else None,
                                 "chunk_edges": report.tuning.chunk_edges if report else None,
                                 "delta": report.delta if report else None},
                   "apply_ms_p50": round(statistics.median(s.apply_ms for s in stats), 4),
                   "engine_ms_p50": round(statistics.median(s.engine_ms for s in stats), 4)},
        "roofline": roofline, "cpu_baseline": cpu, "parity": exact,
    }
    print(json.dumps(out), flush=True)
    del dev, tree, inst
    import gc

    gc.collect()
    dyng_amd.synchronize()
    return 0 if (exact is None or exact["bit_exact"]) else 3


def mosp_inputs(name: str, batches: int):
    """The MOSP workload: an RMAT graph with MOSP_OBJECTIVES weight columns and its batches (the
    extra columns of graph and insertions drawn from seeds of their own)."""
    from dyng_amd import generators as G

    kind, size, m, wmax, cfg, nb, (ni, nd, nr) = WORKLOADS[name]
    nb = max(nb, batches)
    n = 1 << size
    src, dst, w0 = G.rmat(size, m, wmax, G.GRAPH_SEED ^ cfg)
    rng = np.random.default_rng(G.GRAPH_SEED ^ cfg ^ 0x3C)
    w = np.ascontiguousarray(np.stack([w0] + [rng.integers(1, wmax + 1, m, dtype=np.int32)
                                              for _ in range(MOSP_OBJECTIVES - 1)], 1))
    drawn = G.mixed_batches(src, dst, n, nb, ni, nd, nr, wmax, G.BATCH_SEED ^ cfg, rmat_scale=size)
    rng = np.random.default_rng(G.BATCH_SEED ^ cfg ^ 0x3C)
    out = []
    for x in drawn:
        iw = np.ascontiguousarray(np.stack([x[2]] + [rng.integers(1, wmax + 1, len(x[2]), dtype=np.int32)
                                                     for _ in range(MOSP_OBJECTIVES - 1)], 1))
        out.append((x[0], x[1], iw, x[3], x[4]))
    return n, src, dst, w, out, G.max_out_degree_vertex(src, n), ni + nd + nr


def mosp_cpu_baseline(n, src, dst, w, drawn, source, sample: int, threads: int):
    """The oracle on `sample` batches: the batch applied once and the trees repaired by the OpenMP
    engine (update_each), the combined graph, its tree (OpenMP engine) and the path costs."""
    import oracle
    from oracle import mosp as omosp

    row = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(src, minlength=n), out=row[1:])
    g = oracle.Csr(row, dst.astype(np.int32), np.ascontiguousarray(w.T.astype(np.int32)))
    K = w.shape[1]
    ws = [oracle.Workspace() for _ in range(K)]
    trees = [oracle.compute(g, source, k, "omp", threads=threads, ws=ws[k])[0] for k in range(K)]
    parts, results = [], []
    for b in range(sample):
        x = drawn[b]
        t0 = time.perf_counter()
        g, _, cs = oracle.update_many(g, trees, x[0], x[1], x[2], x[3], x[4], engine="omp", threads=threads,
                                      workspaces=ws)
        t1 = time.perf_counter()
        cg, _, _ = omosp.combine_fast([t.parent for t in trees], source, MOSP_PREFERENCES)
        t2 = time.perf_counter()
        ct, _ = oracle.compute(cg, source, 0, "omp", threads=threads)
        t3 = time.perf_counter()
        costs = omosp.path_costs_fast(g, ct.parent, source, K)
        t4 = time.perf_counter()
        repair = sum(sum(c.t_hooks) for c in cs)
        parts.append({"update_each_s": t1 - t0, "repairs_s": repair, "combine_s": t2 - t1, "solve_s": t3 - t2,
                      "costs_s": t4 - t3})
        results.append((ct.dist.copy(), ct.parent.copy(), costs))
        log(f"[bench] cpu mosp batch {b}: update_each {t1 - t0:.2f}s (repairs {repair * 1e3:.1f} ms), combine "
            f"{t2 - t1:.2f}s, solve {t3 - t2:.2f}s, costs {t4 - t3:.2f}s")
    return parts, results


def mosp_bench(args) -> int:
    """The MOSP configuration: mosp.update on RMAT-22 with three objectives."""
    import dyng_amd
    from dyng_amd import mosp

    K, W = args.steps, args.warmup
    t0 = time.perf_counter()
    n, src, dst, w, drawn, source, ops = mosp_inputs(args.workload, K + W)
    log(f"[bench] inputs {args.workload}: n={n} m={len(src)} K={w.shape[1]} batches={len(drawn)}x{ops} in "
        f"{time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    g = dyng_amd.Graph.from_edges(src, dst, w, num_vertices=n)
    r = mosp.compute(g, source, preferences=MOSP_PREFERENCES)
    dyng_amd.synchronize()
    log(f"[bench] device graph + mosp.compute in {time.perf_counter() - t0:.1f}s")
    D = dyng_amd.DeviceArray
    dev = [dyng_amd.EdgeBatch(insert=(D(x[0]), D(x[1]), D(x[2])), delete=(D(x[3]), D(x[4])))
           for x in drawn[:K + W]]
    wall, stats = [], []
    for b in range(W + K):
        dyng_amd.synchronize()
        t1 = time.perf_counter()
        st = mosp.update(g, dev[b], r)
        dt = (time.perf_counter() - t1) * 1e3
        if b >= W:
            wall.append(dt)
            stats.append(st)
    dyng_amd.synchronize()
    p50 = statistics.median(wall)
    value = ops / (p50 / 1e3)
    log(f"[bench] gpu: p50 {p50:.3f} ms/update, {value:.3e} edges/s")
    # the persistent kernel: three repair launches and the combined solve per step
    kbytes, kms = [], []
    for st in stats:
        for o in st.objectives:
            kbytes.append(fused_kernel_bytes(o.device, n, o.batch.deleted_edges + o.batch.ignored_deletions + ops,
                                             ops))
            kms.append(o.device.kernel_ms)
        dels, ins = st.combined_batch
        kbytes.append(fused_kernel_bytes(st.combined_solve, n, dels + ins, ins))
        kms.append(st.combined_solve.kernel_ms)
    achieved = (sum(kbytes) / len(kbytes)) / (statistics.mean(kms) / 1e3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(args.workload),
                "kernel": "sssp_fused_kernel (3 repairs + the combined tree's repair per step)",
                "kernel_ms_mean": round(statistics.mean(kms), 4), "algorithmic_bytes_mean": int(sum(kbytes) / len(kbytes))}
    cpu, exact = None, None
    if not args.no_cpu and args.cpu_sample > 0:
        threads = args.cpu_threads or len(STARTUP_CPUS)
        sample = min(args.cpu_sample, 2)
        parts, results = mosp_cpu_baseline(n, src, dst, w, drawn, source, sample, threads)
        total = [sum(v for k, v in p.items() if k != "repairs_s") for p in parts]
        cpu = {"value": round(ops / statistics.median(total), 1), "unit": "edges/s", "cores": threads,
               "kind": "port", "sample": f"{sample} batches x {ops} ops after compute(); one mosp update = batch "
               "apply + transpose + 3 OpenMP repairs, the combined graph (numpy), its OpenMP solve and the path "
               "costs (numpy)",
               "p50_s_update": round(statistics.median(total), 3),
               "p50_ms_repairs": round(statistics.median(p["repairs_s"] for p in parts) * 1e3, 2)}
        g2 = dyng_amd.Graph.from_edges(src, dst, w, num_vertices=n)
        r2 = mosp.compute(g2, source, preferences=MOSP_PREFERENCES)
        ok = True
        for b in range(sample):
            x = drawn[b]
            mosp.update(g2, dyng_amd.EdgeBatch(insert=(x[0], x[1], x[2]), delete=(x[3], x[4])), r2)
            cd, cp, cc = results[b]
            same = bool(np.array_equal(r2.combined_distances, cd) and np.array_equal(r2.combined_parents, cp)
                        and np.array_equal(r2.path_costs, cc))
            ok = ok and same
            log(f"[bench] parity batch {b}: {'bit-exact' if same else 'MISMATCH'}")
        exact = {"batches_checked": sample, "bit_exact": ok,
                 "against": "oracle (OpenMP repairs, combined graph, OpenMP solve, path costs)"}
        del r2, g2
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "edges/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(sum(wall) / K, 4), "p50_ms": round(p50, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int64", "data": "synthetic",
        "config": {"workload": args.workload, "vertices": int(n), "edges": int(len(src)), "objectives": int(w.shape[1]),
                   "preferences": MOSP_PREFERENCES, "batch_ops": ops, "source": int(source),
                   "apply_ms_p50": round(statistics.median(s.objectives[0].apply_ms for s in stats), 4),
                   "repairs_ms_p50": round(statistics.median(sum(o.engine_ms for o in s.objectives)
                                                             for s in stats), 4),
                   "combined_repair_kernel_ms_p50": round(statistics.median(s.combined_solve.kernel_ms
                                                                            for s in stats), 4),
                   "combined_batch_p50": int(statistics.median(sum(s.combined_batch) for s in stats)),
                   "combined_edges_p50": int(statistics.median(s.combined_edges for s in stats))},
        "roofline": roofline, "cpu_baseline": cpu, "parity": exact,
    }
    print(json.dumps(out), flush=True)
    del dev, r, g
    import gc

    gc.collect()
    dyng_amd.synchronize()
    return 0 if (exact is None or exact["bit_exact"]) else 3


def hyper_bench(args) -> int:
    """hyper_sssp.update on a uniform random hypergraph: the whole update (host bookkeeping of the
    hypergraph and the line graph's edge batch, then the device apply and repair) and its device
    part apart."""
    import dyng_amd
    from dyng_amd import hyper_sssp as hs

    K, W = args.steps, args.warmup
    _, V, H, wmax, _, _, (nd, ni, ne) = WORKLOADS[args.workload]
    from dyng_amd import generators as G

    rng = np.random.default_rng(G.GRAPH_SEED ^ 7)
    sizes = rng.integers(2, 7, H)
    members = rng.integers(0, V, int(sizes.sum()))
    off = np.concatenate([[0], np.cumsum(sizes)])
    lists = [members[off[i]:off[i + 1]].tolist() for i in range(H)]
    weights = rng.integers(1, wmax + 1, H).tolist()
    t0 = time.perf_counter()
    h = dyng_amd.Hypergraph.from_lists(lists, weights, num_vertices=V)
    source = int(np.argmax(np.bincount(members, minlength=V)))
    r = hs.compute(h, source)
    dyng_amd.synchronize()
    lg_edges = r._lg.num_edges
    log(f"[bench] hypergraph V={V} H={H}, line graph {lg_edges} edges, compute in {time.perf_counter() - t0:.1f}s")
    wall, dev_ms, host_ms, lg_changes = [], [], [], []
    for b in range(W + K):
        bound = h.hyperedge_id_bound
        dels = rng.integers(0, bound, nd).tolist()
        ins_sizes = rng.integers(2, 7, ni)
        ins = [rng.integers(0, V, int(k)).tolist() for k in ins_sizes]
        rem = (rng.integers(0, bound, ne).tolist(), rng.integers(0, V, ne).tolist())
        add = (rng.integers(0, bound, ne).tolist(), rng.integers(0, V, ne).tolist())
        batch = dyng_amd.HyperedgeBatch(dels, ins, rng.integers(1, wmax + 1, ni).tolist(),
                                        delete_incidences=rem, insert_incidences=add)
        dyng_amd.synchronize()
        t1 = time.perf_counter()
        st = hs.update(h, batch, r)
        dt = (time.perf_counter() - t1) * 1e3
        if b >= W:
            d = st.device
            wall.append(dt)
            dev_ms.append(d.apply_ms + d.engine_ms)
            host_ms.append(dt - d.apply_ms - d.engine_ms)
            lg_changes.append(d.batch.inserted_edges + d.batch.deleted_edges)
    p50 = statistics.median(wall)
    ops = nd + ni + 2 * ne
    value = ops / (p50 / 1e3)
    again = hs.compute(h, source)
    exact = bool(np.array_equal(again.hyperedge_distances, r.hyperedge_distances)
                 and np.array_equal(again.hyperedge_parents, r.hyperedge_parents))
    log(f"[bench] gpu: p50 {p50:.2f} ms/update (device {statistics.median(dev_ms):.3f} ms), {value:.3e} ops/s, "
        f"updated == compute: {exact}")
    out = {
        "metric": "hyperedge ops/s & p50 ms/update, hyper_sssp on a uniform hypergraph", "value": round(value, 1),
        "unit": "ops/s", "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": round(sum(wall) / K, 4),
        "p50_ms": round(p50, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": args.workload, "vertices": V, "hyperedges": H, "line_graph_edges": int(lg_edges),
                   "batch_ops": ops, "source": source,
                   "device_ms_p50": round(statistics.median(dev_ms), 4),
                   "host_ms_p50": round(statistics.median(host_ms), 4),
                   "line_graph_changes_p50": int(statistics.median(lg_changes)),
                   "device": "sssp.update on the line graph: the device apply of its edge batch and the "
                             "persistent repair",
                   "host": "the hypergraph's bookkeeping and the line graph's edge batch (the C++ container, called from Python)"},
        "parity": {"bit_exact": exact, "against": "hyper_sssp.compute on the final hypergraph"},
    }
    print(json.dumps(out), flush=True)
    return 0 if exact else 3


def main() -> int:
    # the CPU baseline's OpenMP threads, pinned one per core (read when the OpenMP runtime loads,
    # so before the generators' or the oracle's library is opened)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=19)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rmat22_mixed100k", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-sample", type=int, default=5, help="batches timed on the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: every physical core this process may use")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-plain comparison")
    ap.add_argument("--train-batches", type=int, default=200, help="stream workloads: batches the scheduler "
                    "trains on (every schedule timed on each)")
    ap.add_argument("--stream-batches", type=int, default=0, help="stream workloads: evaluated batches")
    ap.add_argument("--computes", type=int, default=3, help="solves from the source timed for the second "
                    "roofline figure")
    ap.add_argument("--compute-only", action="store_true", help="build the graph and run only the timed "
                    "solves from the source (for counter passes over the solve)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported (the path does "
              "not shard across GPUs in this release)", file=sys.stderr)
        return 2

    import dyng_amd
    from dyng_amd import sssp

    if dyng_amd.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        return 1
    if WORKLOADS[args.workload][0] == "stream":
        return stream_bench(args)
    if WORKLOADS[args.workload][0] == "batched":
        return batched_bench(args)
    if WORKLOADS[args.workload][0] == "mosp":
        return mosp_bench(args)
    if WORKLOADS[args.workload][0] == "hyper":
        return hyper_bench(args)
    K, W = args.steps, args.warmup
    n, src, dst, w, batch, source, ops = make_inputs(args.workload, K + W)

    # ---- GPU: graph, initial tree, batches resident in HBM ----
    t0 = time.perf_counter()
    g = dyng_amd.Graph.from_edges(src, dst, w, num_vertices=n)
    tree = sssp.compute(g, source)
    dyng_amd.synchronize()
    log(f"[bench] device graph + compute in {time.perf_counter() - t0:.1f}s")
    D = dyng_amd.DeviceArray

    def device_batch(b):
        x = batch(b, tree.to_numpy()[1] if batch.dynamic else None)
        return dyng_amd.EdgeBatch(insert=(D(x[0]), D(x[1]), D(x[2])), delete=(D(x[3]), D(x[4])))

    def solves(count):
        """`count` solves from the source on the graph as it is: their counters and the
        roofline figure of the persistent kernel over them."""
        cs = []
        for _ in range(count):
            r2 = sssp.compute(g, source)
            cs.append(r2.counters)
            del r2
        cb = [compute_bytes(c, n) for c in cs]
        cms = [c.kernel_ms for c in cs]
        got = (sum(cb) / len(cb)) / (statistics.mean(cms) / 1e3) / 1e9
        log(f"[bench] solve from the source: {statistics.mean(cms):.3f} ms, {got:.1f} GB/s algorithmic, "
            f"{cs[-1].iterations} iterations, {cs[-1].relax_edges} edges scanned, {cs[-1].blocks} workgroups")
        return {"bound": "hbm", "achieved": round(got, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(got / HBM_PEAK_GBS, 4), "traffic": measured_traffic(args.workload + "_compute"),
                "kernel": "sssp_fused_kernel (solve from the source)", "kernel_ms_mean": round(statistics.mean(cms), 4),
                "algorithmic_bytes_mean": int(sum(cb) / len(cb)), "solves": count,
                "relax_edges": int(cs[-1].relax_edges), "iterations": int(cs[-1].iterations),
                "blocks": int(cs[-1].blocks)}

    if args.compute_only:
        res = solves(max(1, args.computes))
        print(json.dumps({"compute_roofline": res}), flush=True)
        del tree, g
        dyng_amd.synchronize()
        return 0

    # mixed batches are uploaded up front; a tree batch is drawn (outside the timed call) against
    # the tree the previous update left
    dev_batches = [None] * (K + W) if batch.dynamic else [device_batch(b) for b in range(K + W)]
    wall, stats = [], []
    for b in range(W + K):
        eb = dev_batches[b] if dev_batches[b] is not None else device_batch(b)
        dyng_amd.synchronize()
        t1 = time.perf_counter()
        st = sssp.update(g, eb, tree)
        dt = (time.perf_counter() - t1) * 1e3
        if b >= W:
            wall.append(dt)
            stats.append(st)
        del eb
    dyng_amd.synchronize()
    total_s = sum(wall) / 1e3
    p50 = statistics.median(wall)
    value = ops / (p50 / 1e3)
    log(f"[bench] gpu: p50 {p50:.3f} ms/batch, mean {total_s / K * 1e3:.3f} ms, {value:.3e} edges/s")

    # ---- roofline of the persistent kernel (HIP events on its stream) ----
    kbytes = [fused_kernel_bytes(s.device, n, s.batch.deleted_edges + s.batch.ignored_deletions + ops,
                                 ops) for s in stats]
    kms = [s.device.kernel_ms for s in stats]
    achieved = (sum(kbytes) / len(kbytes)) / (statistics.mean(kms) / 1e3) / 1e9
    rbytes = [relax_bytes(s.device) for s in stats]
    rms = [s.device.relax_ms for s in stats]
    relax_gbs = (sum(rbytes) / max(sum(rms), 1e-9)) * 1e3 / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(args.workload),
                "kernel": "sssp_fused_kernel", "kernel_ms_mean": round(statistics.mean(kms), 4),
                "algorithmic_bytes_mean": int(sum(kbytes) / len(kbytes))}
    relax_phase = {"achieved_gbs": round(relax_gbs, 1), "frac": round(relax_gbs / HBM_PEAK_GBS, 4),
                   "relax_ms_mean": round(statistics.mean(rms), 4), "bytes_mean": int(sum(rbytes) / len(rbytes)),
                   "teps": round(sum(s.device.relax_edges for s in stats) / max(sum(rms), 1e-9) * 1e3, 1)}

    # ---- where the persistent kernel's time goes: its phases (s_memrealtime stamps) and steps ----
    def mean(f):
        return round(statistics.mean(f(s) for s in stats), 4)

    phases = {"pack_ms": mean(lambda s: s.device.pack_ms), "walk_ms": mean(lambda s: s.device.jump_ms),
              "invalidate_and_pull_ms": mean(lambda s: s.device.step1_ms - s.device.pack_ms - s.device.jump_ms),
              "pull_ms": mean(lambda s: s.device.pull_ms), "relax_ms": mean(lambda s: s.device.relax_ms),
              "unpack_ms": mean(lambda s: s.device.unpack_ms), "kernel_ms": mean(lambda s: s.device.kernel_ms),
              "walk_rounds": mean(lambda s: s.device.jump_rounds), "relax_iterations": mean(lambda s: s.iterations),
              "epochs": mean(lambda s: s.epochs), "invalidated": mean(lambda s: s.invalidated),
              "relax_edges": mean(lambda s: s.device.relax_edges)}
    phases["us_per_walk_round"] = round(1e3 * phases["walk_ms"] / max(phases["walk_rounds"], 1e-9), 2)
    phases["us_per_relax_step"] = round(1e3 * phases["relax_ms"] / max(phases["relax_iterations"]
                                                                        + phases["epochs"], 1e-9), 2)
    log(f"[bench] phases {phases}")
    compute_roofline = solves(args.computes) if args.computes > 0 else None

    # ---- CPU baseline + bit-exactness on the sampled batches ----
    cpu = None
    exact = None
    if not args.no_cpu and args.cpu_sample > 0:
        hc = host_cpus()
        threads = args.cpu_threads or hc["physical_cores"]
        counts = [threads] + ([hc["cores_one_socket"]] if hc["cores_one_socket"] < threads else [])
        summaries, trees, drawn = cpu_baseline(n, src, dst, w, batch, source, args.cpu_sample, counts)
        summary = summaries[threads]
        rep = statistics.median(summary["repair_ms"])
        cpu = {"value": round(ops / (rep / 1e3), 1), "unit": "edges/s", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_sample} batches x {ops} ops after compute(); repair = identify_affected + "
                         f"seed + loop + finalize of the OpenMP engine (median)",
               "samples": args.cpu_sample, "cpu_model": hc["model"], "host": hc,
               "omp_proc_bind": os.environ.get("OMP_PROC_BIND"), "omp_places": os.environ.get("OMP_PLACES"),
               "p50_ms_repair": round(rep, 3), "p50_ms_update_incl_host_apply": round(
                   statistics.median(summary["update_ms"]), 1), "compute_s": round(summary["compute_s"], 2),
               "one_socket": {"threads": hc["cores_one_socket"],
                              "p50_ms_repair": round(statistics.median(summaries[counts[-1]]["repair_ms"]), 3),
                              "same_as_all_cores": len(counts) == 1}}
        # the same batches on a fresh device graph, compared after each
        g2 = dyng_amd.Graph.from_edges(src, dst, w, num_vertices=n)
        t2 = sssp.compute(g2, source)
        ok = True
        for b in range(args.cpu_sample):
            x = drawn[b]
            st = sssp.update(g2, dyng_amd.EdgeBatch(insert=(x[0], x[1], x[2]), delete=(x[3], x[4])), t2)
            d, p = t2.to_numpy()
            cd, cp, inv, aff = trees[b]
            same = bool(np.array_equal(d, cd) and np.array_equal(p, cp) and st.invalidated == inv
                        and st.affected == aff)
            ok = ok and same
            log(f"[bench] parity batch {b}: {'bit-exact' if same else 'MISMATCH'}")
        exact = {"batches_checked": args.cpu_sample, "bit_exact": ok, "against": "OpenMP engine (oracle)"}

    # ---- the learned scheduler against the plain schedule, on the stream it is meant for (a
    # slice of the RMAT-23 stream of uneven batches) ----
    learned = None
    if not args.no_learned:
        count = args.stream_batches or 1000
        sn, sm, ssrc, sdst, sw, ssource, train, evaluate, hindsight, (lo, hi) = stream_inputs(
            "rmat23_learned", count, args.train_batches)
        learned = learned_vs_plain(sn, ssrc, sdst, sw, ssource, train, evaluate(), hindsight)
        learned["workload"] = {"name": "rmat23_learned", "vertices": sn, "edges": sm, "batches": count,
                               "batch_ops": f"log-uniform [{lo}, {hi}]", "delete_fraction": "uniform [0, 1]"}
        del ssrc, sdst, sw, train

    s0 = stats[0].device
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "edges/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(total_s / K * 1e3, 4), "p50_ms": round(p50, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int64", "data": "synthetic",
        "config": {"workload": args.workload, "vertices": int(n), "edges": int(len(src)), "batch_ops": ops,
                   "source": int(source), "engine": "fused persistent (HIP, gfx950)",
                   "apply_ms_p50": round(statistics.median(s.apply_ms for s in stats), 4),
                   "engine_ms_p50": round(statistics.median(s.engine_ms for s in stats), 4),
                   "blocks": int(s0.blocks)},
        "roofline": roofline, "relax_phase": relax_phase, "phases": phases, "compute_roofline": compute_roofline,
        "cpu_baseline": cpu, "parity": exact,
        "learned_vs_plain": learned,
    }
    print(json.dumps(out), flush=True)
    # release every device object before the interpreter (and a profiler's runtime) shuts down
    del dev_batches, tree, g
    if exact is not None:
        del t2, g2
    import gc

    gc.collect()
    dyng_amd.synchronize()
    ok = (exact is None or exact["bit_exact"]) and (learned is None or learned["bit_exact"])
    return 0 if ok else 3


if __name__ == "__main__":
    sys.exit(main())
