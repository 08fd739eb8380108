"""dyng_amd: an MI355X-native implementation of dynG's batch-dynamic SSSP path.

``dyng_amd.sssp.compute`` / ``dyng_amd.sssp.update`` keep the reference's Python API
(python/dyng/sssp.py) on a device-resident graph; ``dyng_amd.update`` applies one batch for
several results (python/dyng/_update.py), ``dyng_amd.mosp`` runs MOSP's combined solve on the same
engine (python/dyng/mosp.py), ``dyng_amd.hyper_sssp`` repairs shortest hyperpaths of a
``dyng_amd.Hypergraph`` on its line graph with the same engine (python/dyng/hyper_sssp.py),
``dyng_amd.io`` reads and writes MOSP's files and
``python -m dyng_amd`` is the ``sssp`` / ``mosp`` command line. The work runs in libdyng_hip.so
(hand-written HIP for gfx950). There is no CPU fallback: without the built library every call
raises NotSupportedError.
"""

from . import batched, hyper_sssp, io, mosp, scheduler, sssp
from ._update import update
from .batch import EdgeBatch
from .device import DeviceArray, device_count, synchronize
from .errors import (DeviceError, DyngError, InvalidArgumentError, NotSupportedError, OutOfMemoryError,
                     StaleResultError)
from .graph import ApplySummary, BatchSemantics, Graph, GraphProperties
from .hypergraph import CapacityError, CapacityPolicy, HyperApplySummary, HyperedgeBatch, Hypergraph, HypergraphProperties
from .resources import Resources

__version__ = "0.1.0"

__all__ = ["update", "sssp", "mosp", "hyper_sssp", "io", "batched", "scheduler", "Graph", "GraphProperties", "BatchSemantics",
           "ApplySummary", "EdgeBatch", "Resources", "Hypergraph", "HypergraphProperties", "HyperedgeBatch",
           "HyperApplySummary", "CapacityPolicy", "CapacityError",
           "DeviceArray", "device_count", "synchronize",
           "DyngError", "InvalidArgumentError", "StaleResultError", "NotSupportedError", "OutOfMemoryError",
           "DeviceError", "available", "__version__"]


def available() -> bool:
    """Whether the HIP library is built and a HIP device is visible (plugin protocol of
    python/plugin/dyng_plugin/__init__.py)."""
    try:
        return device_count() > 0
    except Exception:
        return False
