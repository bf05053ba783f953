"""Device buffers and streams for GPU tests that hand device pointers to the library.

Calls the HIP runtime the device library itself loaded (libamdhip64.so.7, looked up by soname),
so the pointers and streams belong to the same runtime whichever module loaded it first.
"""
from __future__ import annotations

import ctypes

import numpy as np

_hip = None


def hip() -> ctypes.CDLL:
    global _hip
    if _hip is None:
        from gpu_ai_amd.engine import library
        library()  # the device library first: its runtime is the one every call below uses
        lib = ctypes.CDLL("libamdhip64.so.7")
        P, S = ctypes.c_void_p, ctypes.c_size_t
        for name, args in (("hipMalloc", [ctypes.POINTER(P), S]), ("hipFree", [P]),
                           ("hipMemcpy", [P, P, S, ctypes.c_int]),
                           ("hipStreamCreate", [ctypes.POINTER(P)]),
                           ("hipStreamSynchronize", [P]), ("hipStreamDestroy", [P]),
                           ("hipDeviceSynchronize", [])):
            getattr(lib, name).argtypes = args
            getattr(lib, name).restype = ctypes.c_int
        _hip = lib
    return _hip


H2D, D2H = 1, 2


def _ok(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceArray:
    """A device copy of a numpy array (or an uninitialised one of `nbytes`)."""

    def __init__(self, host: np.ndarray | None = None, nbytes: int = 0):
        self.nbytes = host.nbytes if host is not None else nbytes
        self.ptr = ctypes.c_void_p()
        _ok(hip().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")
        if host is not None and self.nbytes:
            self.upload(host)

    def upload(self, host: np.ndarray) -> None:
        h = np.ascontiguousarray(host)
        assert h.nbytes <= self.nbytes
        _ok(hip().hipMemcpy(self.ptr, h.ctypes.data_as(ctypes.c_void_p), h.nbytes, H2D),
            "hipMemcpy H2D")

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        if out.nbytes:
            _ok(hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes, D2H),
                "hipMemcpy D2H")
        return out

    @property
    def address(self) -> int:
        return self.ptr.value

    def free(self) -> None:
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()


class Stream:
    """A HIP stream at the default priority."""

    def __init__(self):
        self.handle = ctypes.c_void_p()
        _ok(hip().hipStreamCreate(ctypes.byref(self.handle)), "hipStreamCreate")

    def synchronize(self) -> None:
        _ok(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self) -> None:
        if self.handle:
            hip().hipStreamDestroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()
