"""Result arenas owned by the library and their DLPack producer side.

A :class:`ResultOwner` wraps ``xtbloom_result_owner_t`` (reference
include/xtbloom/xtbloom.h:1011-1145): one host or ROCm device allocation made
by the library, whose slices a caller binds as output buffers of a compute
call and then hands to any DLPack consumer (``torch.from_dlpack``,
``numpy.from_dlpack``) without copying.  Each export is a managed tensor built
by the library (``xtbloom_result_owner_export_dltensor``) that holds its own
reference on the arena; the PyCapsule destructor below calls its deleter only
when no consumer took it.  Compute calls publish their outputs before they
return (and requests before they report COMPLETE), so an export needs no stream
synchronization.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import library

DLPACK_CPU = 1
DLPACK_ROCM = 10

_NAME_LEGACY = b"dltensor"
_NAME_VERSIONED = b"dltensor_versioned"

_capsule_new = ctypes.pythonapi.PyCapsule_New
_capsule_new.restype = ctypes.py_object
_capsule_new.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
_capsule_is_valid = ctypes.pythonapi.PyCapsule_IsValid
_capsule_is_valid.restype = ctypes.c_int
_capsule_is_valid.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
_capsule_get_pointer = ctypes.pythonapi.PyCapsule_GetPointer
_capsule_get_pointer.restype = ctypes.c_void_p
_capsule_get_pointer.argtypes = [ctypes.c_void_p, ctypes.c_char_p]

_DELETER = ctypes.CFUNCTYPE(None, ctypes.c_void_p)
# offset of the deleter in DLManagedTensor (after the 48-byte DLTensor and
# manager_ctx) and in DLManagedTensorVersioned (after version and manager_ctx)
_DELETER_OFFSET = {_NAME_LEGACY: 56, _NAME_VERSIONED: 16}


def _unconsumed(capsule: int, name: bytes) -> None:
    # a consumer renames the capsule ("used_dltensor..."); an unconsumed one
    # still owns its managed tensor
    if _capsule_is_valid(capsule, name):
        managed = _capsule_get_pointer(capsule, name)
        deleter = ctypes.c_void_p.from_address(managed + _DELETER_OFFSET[name]).value
        if deleter:
            _DELETER(deleter)(managed)


@ctypes.CFUNCTYPE(None, ctypes.c_void_p)
def _destroy_legacy(capsule):  # pragma: no cover - called by the interpreter
    _unconsumed(capsule, _NAME_LEGACY)


@ctypes.CFUNCTYPE(None, ctypes.c_void_p)
def _destroy_versioned(capsule):  # pragma: no cover - called by the interpreter
    _unconsumed(capsule, _NAME_VERSIONED)


_DTYPES = {np.dtype(np.int8): (0, 8), np.dtype(np.int16): (0, 16), np.dtype(np.int32): (0, 32),
           np.dtype(np.int64): (0, 64), np.dtype(np.uint8): (1, 8), np.dtype(np.float32): (2, 32),
           np.dtype(np.float64): (2, 64)}


class ResultOwner:
    """One library-owned result arena (host memory or ROCm device memory)."""

    def __init__(self, size_bytes: int, device_id: int | None = None) -> None:
        lib = library.load()
        opts = library.ResultOwnerOptions()
        _check(lib.xtbloom_result_owner_options_init(ctypes.byref(opts), ctypes.sizeof(opts)))
        if device_id is None:
            opts.memory_space, opts.device_id = library.MEMORY_HOST, -1
        else:
            opts.memory_space, opts.device_id = library.MEMORY_ROCM_DEVICE, int(device_id)
        opts.size_bytes = int(size_bytes)
        handle = ctypes.c_void_p()
        _check(lib.xtbloom_result_owner_create(ctypes.byref(opts), ctypes.byref(handle)))
        self._handle = handle
        buf = library.Buffer()
        _check(lib.xtbloom_result_owner_buffer(handle, ctypes.byref(buf)))
        self.data = int(buf.data)
        self.size_bytes = int(buf.size_bytes)
        self.memory_space = int(buf.memory_space)
        self.device_id = device_id

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._handle

    def buffer(self, byte_offset: int, nbytes: int) -> library.Buffer:
        """An output buffer descriptor for bytes [byte_offset, byte_offset + nbytes)."""
        if byte_offset < 0 or nbytes < 0 or byte_offset + nbytes > self.size_bytes:
            raise ValueError("slice outside the result arena")
        return library.Buffer(self.data + byte_offset, nbytes, self.memory_space, 0)

    def tensor(self, byte_offset: int, dtype, shape) -> "DLPackResultBuffer":
        """A DLPack producer for a compact row-major slice of the arena."""
        return DLPackResultBuffer(self, byte_offset, np.dtype(dtype), tuple(int(x) for x in shape))

    def export(self, byte_offset: int, dtype, shape, version: int) -> int:
        """The address of a fresh managed tensor (DLPack 1.0 when version is 1)."""
        code, bits = _DTYPES[np.dtype(dtype)]
        ext = (ctypes.c_int64 * max(1, len(shape)))(*shape)
        view = library.DLPackView(ctypes.sizeof(library.DLPackView), library.API_VERSION, int(byte_offset),
                                  code, bits, 1, len(shape), 0, ext)
        managed = ctypes.c_void_p()
        _check(library.load().xtbloom_result_owner_export_dltensor(self._handle, ctypes.byref(view),
                                                                   int(version), ctypes.byref(managed)))
        return int(managed.value)

    def close(self) -> None:
        if self._handle:
            library.load().xtbloom_result_owner_release(self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:  # pragma: no cover - interpreter shutdown
            pass


class DLPackResultBuffer:
    """``__dlpack__`` / ``__dlpack_device__`` producer for one arena slice."""

    def __init__(self, owner: ResultOwner, byte_offset: int, dtype: np.dtype, shape: tuple[int, ...]):
        self._owner = owner
        self._offset = byte_offset
        self._dtype = dtype
        self._shape = shape

    def __dlpack_device__(self) -> tuple[int, int]:
        if self._owner.memory_space == library.MEMORY_HOST:
            return (DLPACK_CPU, 0)
        return (DLPACK_ROCM, int(self._owner.device_id))

    def __dlpack__(self, *, stream=None, max_version=None, dl_device=None, copy=None):
        # the bytes are complete when the call that filled them returned
        del stream
        if copy:
            raise BufferError("the result arena is exported without copying")
        if dl_device is not None and tuple(dl_device) != self.__dlpack_device__():
            raise BufferError("the result arena lives on another device")
        versioned = max_version is not None and tuple(max_version)[0] >= 1
        managed = self._owner.export(self._offset, self._dtype, self._shape, 1 if versioned else 0)
        name = _NAME_VERSIONED if versioned else _NAME_LEGACY
        destroy = _destroy_versioned if versioned else _destroy_legacy
        return _capsule_new(managed, name, ctypes.cast(destroy, ctypes.c_void_p))


def _check(status: int) -> None:
    if status != library.STATUS_SUCCESS:
        raise RuntimeError(f"{library.status_string(status)}: {library.last_error()}")
