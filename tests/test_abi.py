"""The C-ABI library loads and exports every entry point include/dyng_hip.h declares (no device
calls: this runs without a GPU)."""

import ctypes
import os
import re

import pytest

from dyng_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "dyng_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dyng_hip_\w+)\s*\(", text)))


def test_header_declares_the_exports_list():
    assert set(declared_functions()) == set(_lib.EXPORTS)


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libdyng_hip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [name for name in declared_functions() if not hasattr(lib, name)]
    assert missing == []


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(Exception, match="not built"):
        _lib.lib()
