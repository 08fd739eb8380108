"""dynG's exception classes (python/dyng/errors.py; cpp/include/dyng/core/error.hpp:28-164)."""


class DyngError(Exception):
    """Base class of every dyng error."""


class InvalidArgumentError(DyngError, ValueError):
    """An argument is invalid (nothing was changed, unless the message says otherwise)."""


class StaleResultError(DyngError):
    """The result does not match the graph's current state, or a failed update poisoned it."""


class NotSupportedError(DyngError):
    """The requested feature, backend or type is not available."""


class OutOfMemoryError(DyngError, MemoryError):
    """Host or device memory ran out."""


class DeviceError(DyngError, RuntimeError):
    """The HIP runtime or a kernel reported an error (the reference's cuda_error)."""


class CapacityError(DyngError):
    """capacity_error (error.hpp): a storage cannot hold the batch."""
