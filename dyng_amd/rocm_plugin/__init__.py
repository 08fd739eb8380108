"""The ROCm plugin of dynG: an import package the reference's backend selection can choose
(python/dyng/_backend.py:1-30; the CUDA plugins' protocol, python/plugin/dyng_plugin/__init__.py).

Registered under the entry point group ``dyng.backends`` (``rocm = dyng_amd.rocm_plugin``), it
offers what the selection asks of a plugin:

- ``available()``: whether it can run here: the HIP library is built and an MI355X (gfx950) is
  visible. It reads the kernel driver's topology (``/sys/class/kfd``) and initializes no GPU
  runtime in the calling process, so a process that only chose its module can still fork workers;
- ``status()``: the reason when it cannot;
- ``__version__``: the dynG version whose bindings it mirrors (0.4.0); the selection loads no plugin
  of another version;
- ``native``: its native module (:mod:`dyng_amd.rocm_plugin._core`), the names of dynG's extension
  module that the sssp path reaches, with the same signatures, over the HIP path of this package.

Nothing here loads the HIP library before ``native`` is read.
"""

from __future__ import annotations

import glob
import importlib
import os
from dataclasses import dataclass
from types import ModuleType

__all__ = ["PLUGIN", "CUDA_MAJOR", "DISTRIBUTION", "Status", "available", "status", "native"]

PLUGIN = "rocm"
#: The selection ranks plugins by CUDA major; this one has none (it ranks after any CUDA plugin
#: of the driver's major, which cannot run on an AMD GPU anyway).
CUDA_MAJOR = None
DISTRIBUTION = "dyng-amd"
__version__ = "0.4.0"

#: gfx950 as the kernel driver's topology writes it (major * 10000 + minor * 100 + stepping).
GFX950 = 90500


@dataclass(frozen=True)
class Status:
    """What :func:`available` found: the GPU targets the driver lists and why it cannot run."""

    driver_version: int | None
    targets: tuple[int, ...]
    reason: str


def _library_path() -> str:
    return os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "libdyng_hip.so")


def _gpu_targets(root: str = "/sys/class/kfd/kfd/topology/nodes") -> tuple[int, ...]:
    """The gfx target versions of the GPU nodes the kernel driver lists (CPU nodes report 0)."""
    out = []
    for props in sorted(glob.glob(os.path.join(root, "*", "properties"))):
        try:
            with open(props) as f:
                for line in f:
                    key, _, value = line.partition(" ")
                    if key == "gfx_target_version" and int(value) > 0:
                        out.append(int(value))
        except (OSError, ValueError):
            continue
    return tuple(out)


def status() -> Status:
    targets = _gpu_targets()
    if not os.path.exists(_library_path()):
        return Status(None, targets, f"the HIP library {_library_path()} is not built (python -m dyng_amd.build)")
    if not os.path.exists("/dev/kfd"):
        return Status(None, targets, "no AMD GPU kernel driver (/dev/kfd is missing)")
    if GFX950 not in targets:
        found = ", ".join(f"gfx{t // 10000}{(t // 100) % 100:x}{t % 100:x}" for t in targets) or "none"
        return Status(None, targets, f"no MI355X (gfx950) is visible (GPU targets: {found})")
    return Status(None, targets, "")


def available() -> bool:
    return status().reason == ""


def __getattr__(name: str) -> ModuleType:
    if name == "native":
        return importlib.import_module("._core", __name__)
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")
