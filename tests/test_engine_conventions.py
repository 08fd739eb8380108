"""The memory-ordering convention of the fused SSSP kernel, checked on its source text.

`sssp_fused_kernel` (dyng_amd/csrc/sssp.hip) separates its phases with a grid barrier that
orders only agent-scope atomic operations (`DYNG_GRID_SYNC`): the barrier issues no cache
write-back or invalidate, so a plain load or store of an array that other workgroups touch
in the same launch may read a stale line or publish nothing.  The kernel therefore touches
every such array only through the agent-scope helpers (`ld`, `st`, `ld_chunk`, `st_chunk`,
`ld_int4`, the `atomic*` family).  The one exception is the distance-only unpack, whose
plain accesses sit next to `DYNG_GRID_SYNC_FENCED`, the barrier variant with a release and
an acquire fence.

The reference's kernel relies on `cudaLaunchCooperativeKernel`'s grid sync, which is a full
fence (dynG sssp/fused.cuh), so it has no such convention to keep.  These tests read the
source text (no compiler, no GPU) and fail on an access that breaks the convention; each
rule is also shown to catch a planted violation.

Rules:
1. An array shared across workgroups (`packed`, `stamp`, `in_far`, `flag`, `ancestor`,
   `cand`, `frontier`, `near_a`, `near_b`, `far_a`, `far_b`, `touched`, `chunks`), or a
   local pointer into one, is subscripted only as `&x[i]`, the first argument of a helper.
   Inside the kernel body a plain access is allowed when the barrier that orders it is
   fenced: a plain store needs the next barrier in text order (or the kernel's end) fenced,
   a plain load the previous one (or the kernel's start).
2. `parents` and `distances` are the run's input tree until the unpack phase: nothing
   writes them before it (plain loads of them are fine anywhere).
Text order stands in for control flow: a store at the end of a loop body whose back edge
reaches an unfenced barrier would pass rule 1.  The kernel has no plain access in a loop
that holds a barrier.
"""

from __future__ import annotations

import pathlib
import re

import pytest

SOURCE = pathlib.Path(__file__).resolve().parents[1] / "dyng_amd" / "csrc" / "sssp.hip"

SHARED = ("packed", "stamp", "in_far", "flag", "ancestor", "cand", "frontier", "near_a", "near_b",
          "far_a", "far_b", "touched", "chunks")
# Local pointers and helper parameters that hold one of the shared arrays.
ALIASES = ("list", "next", "far", "spill", "out", "far_list", "current", "cur", "front", "listed", "q",
           "chunk_in", "chunk_out", "q4", "ch")
TREE = ("parents", "distances")
HELPER = re.compile(r"^(ld|st|ld_chunk|st_chunk|ld_int4|atomic[A-Za-z]+|__hip_atomic_[a-z_]+)$")
KERNEL_START = "sssp_fused_kernel(engine_params p)"
KERNEL_END = "#undef DYNG_GRID_SYNC"
UNPACK = "// ---- F: unpack ----"


def strip_comments(text: str) -> str:
    """Blank out comments and string literals, keeping every offset and line break."""
    out = []
    i, n = 0, len(text)
    while i < n:
        if text.startswith("//", i):
            j = text.find("\n", i)
            j = n if j < 0 else j
            out.append(" " * (j - i))
            i = j
        elif text.startswith("/*", i):
            j = text.find("*/", i + 2)
            j = n if j < 0 else j + 2
            out.append(re.sub(r"[^\n]", " ", text[i:j]))
            i = j
        elif text[i] == '"':
            j = i + 1
            while j < n and text[j] != '"':
                j += 2 if text[j] == "\\" else 1
            out.append('"' + " " * (j - i - 1) + '"')
            i = j + 1
        else:
            out.append(text[i])
            i += 1
    return "".join(out)


def enclosing_call(code: str, pos: int) -> str | None:
    """The name of the call whose argument list holds `pos` (None at statement level)."""
    depth = 0
    i = pos - 1
    while i >= 0:
        c = code[i]
        if c in ")]":
            depth += 1
        elif c == "[":
            depth -= 1
        elif c == "(":
            if depth == 0:
                m = re.search(r"([A-Za-z_]\w*)\s*$", code[:i])
                return m.group(1) if m else None
            depth -= 1
        elif c in ";{}" and depth == 0:
            return None
        i -= 1
    return None


def is_store(code: str, end: int) -> bool:
    """Whether the subscript closing at `end` is assigned to (`x[i] = …`, `x[i] += …`, `x[i]++`)."""
    depth = 1
    i = end
    while depth:
        if code[i] == "[":
            depth += 1
        elif code[i] == "]":
            depth -= 1
        i += 1
    rest = code[i:i + 3].lstrip()
    return bool(re.match(r"(=[^=]|[-+*/|&^]=|\+\+|--)", rest))


def line_of(code: str, pos: int) -> int:
    return code.count("\n", 0, pos) + 1


def violations(text: str) -> list[str]:
    code = strip_comments(text)
    k0 = code.index(KERNEL_START)
    k1 = code.index(KERNEL_END, k0)
    unpack = text.index(UNPACK, k0)
    barriers = [(m.start(), m.group(1) is not None)
                for m in re.finditer(r"DYNG_GRID_SYNC(_FENCED)?\(\)", code[k0:k1])]
    barriers = [(k0 + at, fenced) for at, fenced in barriers]

    def fenced_window(pos: int, store: bool) -> bool:
        if not k0 <= pos < k1:
            return False
        if store:
            after = [f for at, f in barriers if at > pos]
            return after[0] if after else True
        before = [f for at, f in barriers if at < pos]
        return before[-1] if before else True

    found = []
    pattern = r"(?<![\w.>])(p\.)?({})\[".format("|".join(SHARED + ALIASES))
    for m in re.finditer(pattern, code):
        name = m.group(2)
        if m.group(1) is None and name in SHARED:
            continue  # a local of the same name (e.g. `int far_count[2]` is not one of these)
        if m.group(1) is not None and name in ALIASES:
            continue
        start = m.start()
        wrapped = code[:start].rstrip().endswith("&") and bool(HELPER.match(enclosing_call(code, start) or ""))
        if wrapped:
            continue
        if fenced_window(start, is_store(code, m.end())):
            continue
        found.append(f"line {line_of(code, start)}: plain access to {m.group(0)[:-1]}")
    for m in re.finditer(r"p\.({})\[".format("|".join(TREE)), code):
        if m.start() >= unpack:
            continue
        store = is_store(code, m.end())
        atomic = code[:m.start()].rstrip().endswith("&") and (enclosing_call(code, m.start()) or "").startswith(
            ("atomic", "st", "__hip_atomic"))
        if store or atomic:
            found.append(f"line {line_of(code, m.start())}: {m.group(0)[:-1]} written before the unpack phase")
    return found


@pytest.fixture(scope="module")
def source() -> str:
    return SOURCE.read_text()


def test_kernel_keeps_the_ordering_convention(source):
    assert violations(source) == []


def test_the_distance_only_unpack_is_the_only_plain_section(source):
    """The plain accesses the rules let through all sit in the distance-only unpack."""
    code = strip_comments(source)
    k0 = code.index(KERNEL_START)
    fenced = [line_of(code, m.start()) for m in re.finditer(r"DYNG_GRID_SYNC_FENCED\(\)", code[k0:])]
    assert fenced, "the distance-only unpack lost its fenced barriers"
    plain = []
    for m in re.finditer(r"(?<![\w.>&])p\.({})\[".format("|".join(SHARED)), code):
        plain.append(line_of(code, m.start()))
    assert plain and all(line > min(fenced) - 12 for line in plain), plain


def _plant(source: str, anchor: str, replacement: str) -> str:
    assert source.count(anchor) >= 1, anchor
    return source.replace(anchor, replacement, 1)


@pytest.mark.parametrize("anchor, replacement", [
    # a plain load of a word another workgroup may have lowered
    ("seen[j] = live[j] ? ld(&p.packed[w[j]]) : 0;", "seen[j] = live[j] ? p.packed[w[j]] : 0;"),
    # a plain store into a list other workgroups read after the barrier
    ("if (pred) st(&list[slot], v);", "if (pred) list[slot] = v;"),
    # a plain load through a local alias of a shared list
    ("v = ld(&cur[idx]);", "v = cur[idx];"),
    # a shared flag written through a helper that is not agent-scope
    ("if (near) st(&p.in_far[v], 0);", "if (near) write_plain(&p.in_far[v], 0);"),
])
def test_a_planted_plain_access_is_caught(source, anchor, replacement):
    assert violations(_plant(source, anchor, replacement))


def test_a_fenced_barrier_made_unfenced_is_caught(source):
    """The distance-only unpack's plain stores need its barrier to fence."""
    code = strip_comments(source)
    at = code.index("DYNG_GRID_SYNC_FENCED()", code.index(KERNEL_START))
    planted = source[:at] + "DYNG_GRID_SYNC()       " + source[at + len("DYNG_GRID_SYNC_FENCED()"):]
    assert violations(planted)


def test_a_tree_write_before_unpack_is_caught(source):
    planted = _plant(source, "for (int v = tid; v < n; v += threads) st(&p.ancestor[v], p.parents[v]);",
                     "for (int v = tid; v < n; v += threads) p.parents[v] = -1;")
    assert any("before the unpack" in v for v in violations(planted))
