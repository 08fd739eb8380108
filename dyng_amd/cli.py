"""The command line of the HIP path: ``python -m dyng_amd <sssp|mosp> <compute|update>``
(python/dyng/cli/__init__.py, python/dyng/cli/_algorithms.py:184-520).

    python -m dyng_amd sssp compute --graph roadNet-CA_ --out init                   # = mospPrep init
    python -m dyng_amd sssp update --graph roadNet-CA_ --changes b1 --init init --out updated
    python -m dyng_amd mosp update --graph roadNet-CA_ --changes b1 --init init \\
        --preferences 4,1,4 --out updated                                           # = mosp --pref 4,1,4

The graph is MOSP's text CSR (``<prefix>RowPtr.txt``, ``ColInd.txt``, ``Values.txt``), the batch
MOSP's ``insert.txt`` / ``delete.txt`` (read with MOSP's accept/reject rules), and the outputs
MOSP's files, byte for byte: ``<out>/obj<k>/distances<Suffix>.txt`` and ``SSSPTree<Suffix>.txt``
(Suffix ``Original`` for compute, ``Updated`` for update), and for mosp
``<out>/combinedGraph/distancesCsr.txt``, ``SSSPTreeCsr.txt`` and ``mospCosts.txt`` (the costs of
every weight column of the graph along the MOSP tree). Report lines go to standard output
(``--quiet`` drops them). Exit status: 0 on success, 1 when the work fails
(``dyng: error: ...``), 2 for a usage error.
"""

from __future__ import annotations

import argparse
import ctypes as C
import sys
import time
from pathlib import Path

import numpy as np

__all__ = ["build_parser", "main"]


class CliError(Exception):
    """A failure reported as ``dyng: error: <message>`` with exit status 1."""


class UsageError(CliError):
    """A command-line mistake: exit status 2."""


def _info(args, text: str) -> None:
    if not args.quiet:
        print(text, flush=True)


def _fmt_ms(ms: float) -> str:
    return f"{ms:.6g}"


def _preferences(text: str | None) -> list[int]:
    if not text:
        return []
    try:
        return [int(x) for x in text.split(",")]
    except ValueError:
        raise UsageError(f"--preferences: expected comma-separated integers, got {text!r}") from None


def _add_common(p: argparse.ArgumentParser, *, update: bool, mosp: bool) -> None:
    g = p.add_argument_group("graph input")
    g.add_argument("--graph", "--input", dest="graph", required=True, metavar="PREFIX",
                   help="MOSP's text CSR prefix (<prefix>RowPtr.txt, ColInd.txt, Values.txt)")
    g.add_argument("--format", choices=("auto", "csr"), default="auto", help="the graph format (csr)")
    g.add_argument("--num-weights", type=int, default=None, metavar="K",
                   help="the objectives of a graph without edges (as mospPrep -k)")
    g.add_argument("--properties", choices=("mosp_compatible", "default"), default="mosp_compatible",
                   help="graph properties preset (default: mosp_compatible)")
    p.add_argument("--source", type=int, default=0, metavar="S", help="source vertex (default 0)")
    o = p.add_argument_group("mosp options" if mosp else "sssp options")
    o.add_argument("--delta", type=int, default=0, help="near-far bucket width (0: automatic)")
    o.add_argument("--cuda-engine", "--engine", dest="engine", choices=("automatic", "fused"), default="automatic",
                   help="the engine (the fused persistent kernel)")
    if mosp:
        o.add_argument("--preferences", default=None, metavar="P1,P2,..",
                       help="one preference per objective (default: all 1); MOSP's --pref")
        o.add_argument("--num-objectives", type=int, default=0, metavar="K",
                       help="the first K weight columns are the objectives (0: every column); MOSP's -k")
        o.add_argument("--no-path-costs", action="store_true", help="do not compute or write mospCosts.txt")
    else:
        o.add_argument("--objective", type=int, default=None, metavar="K",
                       help="one objective (default: every weight column, one tree each)")
    if update:
        o.add_argument("--no-validate-inputs", action="store_true", help="--init: skip the checks of the trees")
        b = p.add_argument_group("batch (MOSP's insert.txt / delete.txt)")
        b.add_argument("--changes", metavar="DIR", default=None,
                       help="the directory of insert.txt ('u v w1 .. wK' per line) and delete.txt ('u v')")
        b.add_argument("--insert", metavar="FILE", default=None, help="the insertion file")
        b.add_argument("--delete", metavar="FILE", default=None, help="the deletion file")
        p.add_argument("--init", metavar="DIR", default=None,
                       help="the initial trees <init>/obj<k>/distancesOriginal.txt and SSSPTreeOriginal.txt "
                            "(default: computed)")
        p.add_argument("--canonicalize", action="store_true",
                       help="--init: apply the lowest-id tie rule to the initial trees")
        p.add_argument("--write-graph", metavar="PREFIX", default=None,
                       help="also write the updated graph as <PREFIX>{RowPtr,ColInd,Values}.txt")
    p.add_argument("--out", required=True, metavar="DIR", help="output directory")
    p.add_argument("--quiet", action="store_true", help="no report lines on standard output")
    p.add_argument("--device", type=int, default=0, metavar="D", help="HIP device")


def build_parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(prog="python -m dyng_amd", description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="algorithm", metavar="ALGORITHM", required=True)
    for name, mosp in (("sssp", False), ("mosp", True)):
        p = sub.add_parser(name, help=("dynamic multi-objective shortest paths" if mosp
                                       else "dynamic single-source shortest paths"))
        verbs = p.add_subparsers(dest="verb", metavar="VERB", required=True)
        for verb in ("compute", "update"):
            v = verbs.add_parser(verb)
            _add_common(v, update=verb == "update", mosp=mosp)
            v.set_defaults(func=globals()[f"_run_{name}_{verb}"])
    return ap


# ------------------------------------------------------------------------------------------------


def _resources(args):
    import dyng_amd

    return dyng_amd.Resources.hip(args.device)


def _load_graph(args):
    from . import io

    prefix = args.graph
    if not Path(f"{prefix}RowPtr.txt").is_file():
        raise CliError(f"{prefix}RowPtr.txt: no such file (--graph is MOSP's CSR prefix)")
    k = 0
    if args.num_weights:
        values = Path(f"{prefix}Values.txt")
        try:
            empty = not values.read_text(encoding="ascii", errors="replace").strip()
        except OSError:
            empty = False
        k = args.num_weights if empty else 0
    g = io.read_csr_triplet(prefix, num_weights=k, properties=args.properties, resources=_resources(args))
    if not 0 <= args.source < g.num_vertices:
        raise CliError(f"--source {args.source} is out of range: the graph has {g.num_vertices} vertices "
                       f"(0 to {g.num_vertices - 1})")
    return g


def _batch_paths(args):
    if args.changes is not None:
        if args.insert is not None or args.delete is not None:
            raise UsageError("give either --changes or --insert/--delete")
        return Path(args.changes) / "insert.txt", Path(args.changes) / "delete.txt"
    if args.insert is None and args.delete is None:
        raise UsageError("a batch is required: --changes DIR or --insert FILE --delete FILE")
    if args.insert is None or args.delete is None:
        raise UsageError("--insert and --delete go together")
    return Path(args.insert), Path(args.delete)


def _read_batch(args, g):
    from . import io

    ipath, dpath = _batch_paths(args)
    return io.read_legacy_batch(ipath, dpath, num_weights=g.num_weights, num_vertices=g.num_vertices,
                                mosp_lenient=True)


def _objective_dir(out, k: int) -> Path:
    d = Path(out) / f"obj{k}"
    d.mkdir(parents=True, exist_ok=True)
    return d


def _write_tree(out, k: int, suffix: str, distances, parents) -> None:
    from . import io

    d = _objective_dir(out, k)
    io.write_distances(d / f"distances{suffix}.txt", distances)
    io.write_parents(d / f"SSSPTree{suffix}.txt", parents)


def _read_init(args, g, k: int):
    from . import io

    d = Path(args.init) / f"obj{k}"
    return (io.read_distances(d / "distancesOriginal.txt", g.num_vertices),
            io.read_parents(d / "SSSPTreeOriginal.txt", g.num_vertices))


def _run_sssp_compute(args) -> int:
    from . import sssp

    g = _load_graph(args)
    objectives = [args.objective] if args.objective is not None else list(range(max(g.num_weights, 1)))
    for k in objectives:
        t0 = time.perf_counter()
        tree = sssp.compute(g, args.source, objective=k, delta=args.delta, cuda_engine=args.engine)
        d, p = tree.to_numpy()
        ms = (time.perf_counter() - t0) * 1e3
        _write_tree(args.out, k, "Original", d, p)
        reachable = int(np.count_nonzero(d < sssp.INFINITE_DISTANCE))
        _info(args, f"obj{k}: compute {_fmt_ms(ms)} ms (reachable {reachable} of {g.num_vertices})")
    return 0


def _run_sssp_update(args) -> int:
    import dyng_amd
    from . import io, sssp

    g = _load_graph(args)
    batch = _read_batch(args, g)
    objectives = [args.objective] if args.objective is not None else list(range(max(g.num_weights, 1)))
    trees = []
    for k in objectives:
        opt = dict(objective=k, delta=args.delta, cuda_engine=args.engine)
        if args.init is None:
            trees.append(sssp.compute(g, args.source, **opt))
        else:
            d, p = _read_init(args, g, k)
            trees.append(sssp.Result.from_arrays(g, args.source, d, p, canonicalize=args.canonicalize,
                                                 validate_inputs=not args.no_validate_inputs, **opt))
    t0 = time.perf_counter()
    stats = dyng_amd.update(g, batch, *trees)
    ms = (time.perf_counter() - t0) * 1e3
    for k, tree, st in zip(objectives, trees, stats):
        d, p = tree.to_numpy()
        _write_tree(args.out, k, "Updated", d, p)
        _info(args, f"obj{k}: invalidated {st.invalidated}, affected {st.affected}, iterations "
                    f"{st.iterations}, engine {st.engine_used}")
    b = stats[0].batch
    _info(args, f"update {_fmt_ms(ms)} ms ({len(trees)} tree(s); batch: inserted {b.inserted_edges}, updated "
                f"{b.updated_edges}, deleted {b.deleted_edges}, ignored deletions {b.ignored_deletions})")
    if args.write_graph is not None:
        io.write_csr_triplet(args.write_graph, g)
    return 0


def _mosp_options(args, g):
    from . import mosp

    k = args.num_objectives or g.num_weights
    if not 0 <= args.num_objectives <= g.num_weights:
        raise CliError(f"--num-objectives {args.num_objectives} is out of range: the graph has "
                       f"{g.num_weights} weight column(s)")
    opt = mosp.Options(preferences=_preferences(args.preferences), delta=args.delta, cuda_engine=args.engine,
                       compute_path_costs=not args.no_path_costs, num_objectives=args.num_objectives,
                       validate_inputs=not getattr(args, "no_validate_inputs", False))
    return opt, k


def _write_mosp(out, suffix: str, g, paths) -> None:
    """The K trees and the combined outputs; mospCosts.txt covers every weight column of the
    graph (with fewer objectives the other columns are summed along the same MOSP tree)."""
    from . import io, mosp

    for k in range(paths.num_objectives):
        d, p = paths.tree(k).to_numpy()
        _write_tree(out, k, suffix, d, p)
    comb = Path(out) / "combinedGraph"
    comb.mkdir(parents=True, exist_ok=True)
    io.write_distances(comb / "distancesCsr.txt", paths.combined_distances)
    io.write_parents(comb / "SSSPTreeCsr.txt", paths.combined_parents)
    if not paths.options.compute_path_costs:
        return
    if paths.num_objectives == g.num_weights:
        costs = paths.path_costs
    else:
        costs = mosp.path_costs_over(g, paths, g.num_weights)
    io.write_path_costs(comb / "mospCosts.txt", costs)


def _combined_line(paths, st) -> str:
    from . import sssp

    reachable = int(np.count_nonzero(paths.combined_distances < sssp.INFINITE_DISTANCE))
    tail = "" if st is None else f", {st.combined_edges} edges, affected {st.affected}"
    return f"combined: L={paths.preference_scale}, reachable {reachable} of {len(paths.combined_parents)}{tail}"


def _run_mosp_compute(args) -> int:
    from . import mosp

    g = _load_graph(args)
    opt, k = _mosp_options(args, g)
    t0 = time.perf_counter()
    paths = mosp.compute(g, args.source, options=opt)
    ms = (time.perf_counter() - t0) * 1e3
    _write_mosp(args.out, "Original", g, paths)
    _info(args, f"compute {_fmt_ms(ms)} ms ({k} objective(s))")
    _info(args, _combined_line(paths, None))
    return 0


def _run_mosp_update(args) -> int:
    from . import io, mosp

    g = _load_graph(args)
    batch = _read_batch(args, g)
    opt, k = _mosp_options(args, g)
    if args.init is None:
        paths = mosp.compute(g, args.source, options=opt)
    else:
        dist, par = zip(*(_read_init(args, g, i) for i in range(k)))
        paths = mosp.Result.from_arrays(g, args.source, list(dist), list(par), canonicalize=args.canonicalize,
                                        options=opt)
    t0 = time.perf_counter()
    st = mosp.update(g, batch, paths)
    ms = (time.perf_counter() - t0) * 1e3
    _write_mosp(args.out, "Updated", g, paths)
    for i, s in enumerate(st.objectives):
        _info(args, f"obj{i}: invalidated {s.invalidated}, affected {s.affected}, iterations {s.iterations}, "
                    f"engine {s.engine_used}")
    _info(args, _combined_line(paths, st))
    b = st.batch
    _info(args, f"update {_fmt_ms(ms)} ms ({k} objective(s); batch: inserted {b.inserted_edges}, updated "
                f"{b.updated_edges}, deleted {b.deleted_edges}, ignored deletions {b.ignored_deletions})")
    if args.write_graph is not None:
        io.write_csr_triplet(args.write_graph, g)
    return 0


def main(argv=None, stdout=None) -> int:
    """Run the command line; returns the exit status."""
    from .errors import DyngError

    parser = build_parser()
    try:
        args = parser.parse_args(argv)
    except SystemExit as e:
        return int(e.code or 0)
    try:
        return int(args.func(args))
    except UsageError as e:
        print(f"dyng: error: {e}", file=sys.stderr, flush=True)
        return 2
    except (CliError, DyngError, OSError) as e:
        text = str(e)
        print(f"dyng: error: {text[len('dyng: '):] if text.startswith('dyng: ') else text}", file=sys.stderr,
              flush=True)
        return 1
