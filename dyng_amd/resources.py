"""dyng_amd.Resources: the execution resources of the HIP backend (dyng::resources::hip(),
mirroring resources::cuda(int device, stream_ref, int host_threads), core/resources.hpp:135)."""

from __future__ import annotations

from dataclasses import dataclass


@dataclass(frozen=True)
class Resources:
    backend: str = "hip"
    device: int = 0
    stream: object = None  # None: the default stream; else a raw hipStream_t handle (int)

    @staticmethod
    def hip(device: int = 0, stream: object = None) -> "Resources":
        return Resources("hip", int(device), stream)

    @property
    def stream_handle(self):
        s = self.stream
        if s is None:
            return None
        if isinstance(s, int):
            return s
        handle = getattr(s, "cuda_stream", None)
        if handle is None:
            raise TypeError("Resources.stream must be None or an int hipStream_t handle")
        return int(handle)


_DEFAULT = Resources.hip()


def resolve(res: Resources | None, fallback: Resources | None = None) -> Resources:
    if res is not None:
        if not isinstance(res, Resources):
            raise TypeError("resources must be a dyng_amd.Resources")
        return res
    return fallback or _DEFAULT
