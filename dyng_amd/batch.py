"""dyng_amd.EdgeBatch: insertions (upserts), deletions and weight changes
(python/dyng/batch.py; C++ edge_batch_view, cpp/include/dyng/graph/edge_batch.hpp:38-76).

Arrays may live in host memory (numpy, lists) or in device memory (dyng_amd.DeviceArray, int32);
a device batch is read in place by the device apply.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .errors import InvalidArgumentError


def _is_device(x) -> bool:
    from .device import DeviceArray

    return isinstance(x, DeviceArray)


def _pair(value, what: str, with_weights: bool):
    if value is None:
        return None
    parts = tuple(value)
    want = 3 if with_weights else 2
    if len(parts) != want:
        raise InvalidArgumentError(f"EdgeBatch: {what} must be a tuple of {want} arrays")
    return parts


class EdgeBatch:
    """A batch: `insert=(src, dst, weights)` and `delete=(src, dst)`."""

    __slots__ = ("_ins", "_del", "_device")

    def __init__(self, insert=None, delete=None):
        ins = _pair(insert, "insert", True)
        dele = _pair(delete, "delete", False)
        arrays = [a for p in (ins or ()) + (dele or ()) for a in (p,)]
        dev = [_is_device(a) for a in arrays]
        if any(dev) and not all(dev):
            raise InvalidArgumentError("EdgeBatch: mix of host and device arrays")
        self._device = bool(dev) and all(dev)
        if self._device:
            self._ins = self._check_device(ins, True) if ins else None
            self._del = self._check_device(dele, False) if dele else None
        else:
            self._ins = self._host(ins, True) if ins else None
            self._del = self._host(dele, False) if dele else None

    @staticmethod
    def _host(parts, with_weights):
        ids = []
        for a in parts[:2]:
            x = np.asarray(a)
            if x.size and not np.issubdtype(x.dtype, np.integer):
                raise InvalidArgumentError("EdgeBatch: ids must be integers")
            x = x.astype(np.int64).reshape(-1)
            if x.size and (x.max() >= 2**31 - 1 or x.min() < -(2**31)):
                raise InvalidArgumentError("EdgeBatch: ids must fit int32")
            ids.append(np.ascontiguousarray(x.astype(np.int32)))
        if len(ids[0]) != len(ids[1]):
            raise InvalidArgumentError(f"EdgeBatch: the batch has {len(ids[0])} sources but {len(ids[1])} destinations")
        if not with_weights:
            return tuple(ids)
        w = np.asarray(parts[2])
        if w.size and not np.issubdtype(w.dtype, np.integer):
            raise InvalidArgumentError("EdgeBatch: sssp needs integer weights")
        w = w.astype(np.int64)
        if w.ndim == 1:
            w = w.reshape(-1, 1)
        if w.shape[0] != len(ids[0]):
            raise InvalidArgumentError("EdgeBatch: one weight row per insertion is required")
        if w.size and (w.max() > 2**31 - 1 or w.min() < -(2**31)):
            raise InvalidArgumentError("EdgeBatch: weights must fit int32")
        return ids[0], ids[1], np.ascontiguousarray(w.astype(np.int32))

    @classmethod
    def _from_library(cls, insert, delete) -> "EdgeBatch":
        """A host batch from contiguous int32 arrays the library produced: (src, dst, weights) and
        (src, dst), one weight per insertion. It skips the checks and conversions of the
        constructor; the fields are the ones the constructor would build from the same arrays."""
        b = cls.__new__(cls)
        b._device = False
        b._ins = (insert[0], insert[1], insert[2].reshape(-1, 1))
        b._del = (delete[0], delete[1])
        return b

    @staticmethod
    def _check_device(parts, with_weights):
        out = list(parts)
        for a in out:
            if a.dtype != np.int32:
                raise InvalidArgumentError("EdgeBatch: device arrays must be int32")
        if out[0].size != out[1].size:
            raise InvalidArgumentError("EdgeBatch: sources and destinations differ in length")
        if with_weights:
            w = out[2]
            if w.size % max(out[0].size, 1) != 0 or (out[0].size == 0 and w.size != 0):
                raise InvalidArgumentError("EdgeBatch: one weight row per insertion is required")
        return tuple(out)

    @property
    def on_device(self) -> bool:
        return self._device

    @property
    def num_insertions(self) -> int:
        if self._ins is None:
            return 0
        return int(self._ins[0].size) if self._device else len(self._ins[0])

    @property
    def num_deletions(self) -> int:
        if self._del is None:
            return 0
        return int(self._del[0].size) if self._device else len(self._del[0])

    @property
    def num_weights(self) -> int:
        if self._ins is None:
            return 0
        if self._device:
            return int(self._ins[2].size // max(self._ins[0].size, 1))
        return int(self._ins[2].shape[1])

    def empty(self) -> bool:
        return self.num_insertions == 0 and self.num_deletions == 0

    def _part(self, which, i):
        part = self._ins if which == "insert" else self._del
        if part is None:
            if which == "insert" and i == 2:
                return np.zeros((0, 0), np.int32)
            return np.zeros(0, np.int32)
        return part[i]

    @property
    def insert_src(self):
        """Insertion sources (int32; a DeviceArray for a device batch)."""
        return self._part("insert", 0)

    @property
    def insert_dst(self):
        return self._part("insert", 1)

    @property
    def insert_weights(self):
        """(num_insertions, K) int32 (a DeviceArray for a device batch)."""
        return self._part("insert", 2)

    @property
    def delete_src(self):
        return self._part("delete", 0)

    @property
    def delete_dst(self):
        return self._part("delete", 1)

    def _native(self, graph_weights: int):
        """The C struct and the objects that must outlive the call."""
        ni, nd = self.num_insertions, self.num_deletions
        keep = []

        def ptr(a):
            if a is None:
                return None
            keep.append(a)
            return a.ptr if self._device else a.__array_interface__["data"][0]

        ins = self._ins or (None, None, None)
        dele = self._del or (None, None)
        K = self.num_weights if ni else graph_weights
        if ni and K != graph_weights:
            raise InvalidArgumentError(
                f"sssp::update: the batch has {K} weight(s) per insertion but the graph has {graph_weights}")
        nb = _lib.Batch(ptr(ins[0]) if ni else None, ptr(ins[1]) if ni else None, ptr(ins[2]) if ni else None,
                        ni, ptr(dele[0]) if nd else None, ptr(dele[1]) if nd else None, nd, K,
                        int(self._device))
        return nb, keep

    def __repr__(self) -> str:
        where = "device" if self._device else "host"
        return f"dyng_amd.EdgeBatch(insertions={self.num_insertions}, deletions={self.num_deletions}, {where})"
