"""dyng_amd.update(graph, batch, r1, r2, ...): one batch, several results (C++ ``dyng::update`` /
``update_each``, cpp/include/dyng/update.hpp:263-364; Python python/dyng/_update.py).

The batch is applied to the device graph once and every result is brought up to date on the new
graph: each sssp result is repaired by the persistent engine (any objective, any source), each
mosp result's K trees likewise, then its combined graph, MOSP tree and path costs are rebuilt.
Every check (a stale, duplicated or foreign result, an invalid batch) runs before the batch is
applied, so nothing changes when one fails; after the commit a result whose repair fails is
poisoned, the others are still updated, and the first error is raised.
"""

from __future__ import annotations

import ctypes as C
from typing import Any

from . import _lib
from .batch import EdgeBatch
from .errors import InvalidArgumentError
from .graph import Graph
from .resources import Resources, resolve

__all__ = ["update"]


def update_trees(graph: Graph, batch: EdgeBatch, trees: list, resources: Resources | None = None) -> list:
    """Apply ``batch`` once and repair every sssp result in ``trees`` (one C call); returns one
    :class:`dyng_amd.sssp.Stats` per tree (the apply time in the first)."""
    from . import sssp

    res = resolve(resources, graph.resources)
    nb, keep = batch._native(graph.num_weights)
    handles = (C.c_void_p * len(trees))(*[t._handle for t in trees])
    out = (_lib.Stats * len(trees))()
    _lib.check(_lib.lib().dyng_hip_api_update_each(graph._handle, C.byref(nb), handles, len(trees),
                                                   res.stream_handle, out))
    del keep
    return [sssp.Stats._from_native(out[i]) for i in range(len(trees))]


def update(graph: Graph, batch: EdgeBatch, *results: Any, resources: Resources | None = None) -> tuple:
    """Apply ``batch`` to ``graph`` once and bring every result (``dyng_amd.sssp.Result``,
    ``dyng_amd.mosp.Result``) up to date; one stats object per result, in order.

    Raises StaleResultError when a result does not match the graph, InvalidArgumentError for no
    result, a result passed twice or an invalid batch (nothing changed in either case)."""
    from . import mosp, sssp

    if not isinstance(graph, Graph):
        raise TypeError("dyng_amd.update: expected a dyng_amd.Graph")
    if not isinstance(batch, EdgeBatch):
        raise TypeError("dyng_amd.update: batch must be a dyng_amd.EdgeBatch")
    if not results:
        raise InvalidArgumentError("dyng_amd.update: no result (applying a batch without updating any result "
                                   "would leave every result of the graph stale; use Graph.apply)")
    for i, r in enumerate(results):
        if not isinstance(r, (sssp.Result, mosp.Result)):
            raise TypeError(f"dyng_amd.update: result {i} is a {type(r).__name__}; expected dyng_amd.sssp.Result "
                            "or dyng_amd.mosp.Result")
        if any(r is q for q in results[:i]):
            raise InvalidArgumentError(f"dyng_amd.update: result {i} is passed twice")
    # every tree to repair, in order: an sssp result's own, a mosp result's K trees
    trees, spans = [], []
    for r in results:
        own = [r] if isinstance(r, sssp.Result) else list(r._trees)
        spans.append((len(trees), len(own)))
        trees.extend(own)
    per = update_trees(graph, batch, trees, resources)
    out = []
    for r, (a, count) in zip(results, spans):
        if isinstance(r, sssp.Result):
            out.append(per[a])
        else:
            out.append(mosp._finish_update(graph, r, tuple(per[a:a + count])))
    return tuple(out)
