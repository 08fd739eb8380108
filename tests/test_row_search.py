"""The row searches of the device apply and the engine (dyng_amd/csrc/row_search.hpp), built for
the host with g++ and checked against std::lower_bound / std::upper_bound on masked neighbour
ids (tests/native/row_search_check.cpp): rows with runs of parallel edges and dead marks at every
offset within a 16-byte word, every sub-range of short rows, random ranges of long ones. A second
build with AddressSanitizer shows that the last round's aligned 16-byte loads stay within the
words that hold the row."""

from __future__ import annotations

import pathlib
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("sanitize", [False, True], ids=["plain", "asan"])
def test_row_searches_match_the_standard_library(tmp_path, sanitize):
    exe = tmp_path / "row_search_check"
    extra = ["-g", "-fsanitize=address", "-fno-omit-frame-pointer"] if sanitize else []
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Wno-unknown-pragmas", *extra, "-o", str(exe),
                    str(ROOT / "tests" / "native" / "row_search_check.cpp")], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok ")
