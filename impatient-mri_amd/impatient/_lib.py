"""ctypes binding of libimpatient.so (the C ABI declared in include/impatient.h).

The library is built in-tree (``make -C impatient-mri_amd``) and loaded from
``impatient-mri_amd/lib/libimpatient.so``.  Loading fails loudly when it is
missing: there is no CPU fallback for any operator.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_ROOT = Path(__file__).resolve().parent.parent          # impatient-mri_amd/
LIB_PATH = PKG_ROOT / "lib" / "libimpatient.so"
REPO_ROOT = PKG_ROOT.parent
HEADER_PATH = REPO_ROOT / "include" / "impatient.h"

IMP_OK = 0
STATUS_NAMES = {0: "IMP_OK", -1: "IMP_EINVAL", -2: "IMP_EHIP", -3: "IMP_EIO",
                -4: "IMP_ENOMEM", -5: "IMP_EFORMAT", -6: "IMP_ENODEV",
                -7: "IMP_EBREAKDOWN", -8: "IMP_EUNSUPPORTED"}

IMP_Q_CUMULATIVE = 0x1
IMP_Q_DIRECT = 0x2
IMP_CG_BEST_ITERATE = 0x1
IMP_CG_USE_X0 = 0x2
IMP_CG_TRUE_RESIDUAL = 0x4
IMP_CG_PRECOND = 0x8
IMP_CG_FALLBACK_ZERO = 0x10
IMP_CG_GRAPH = 0x20
IMP_COMPAT_Q_CUMULATIVE = 0x1


class ImpError(RuntimeError):
    def __init__(self, status: int, where: str, msg: str):
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class CGStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int),
                ("best_iteration", C.c_int), ("rel_residual", C.c_double),
                ("best_rel_residual", C.c_double), ("true_rel_residual", C.c_double),
                ("seconds", C.c_double), ("start_used", C.c_int),
                ("start_rel_residual", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class NudftJob(C.Structure):
    """imp_nudft_job (device pointers as integers)."""
    _fields_ = [("num_k", C.c_int), ("num_x", C.c_int), ("ncoils", C.c_int)] + [
        (n, C.c_void_p) for n in ("kx", "ky", "kz", "t", "x", "y", "z", "fm", "phi_r", "phi_i",
                                  "d_r", "d_i", "s_r", "s_i", "out_r", "out_i")]


class FileInfo(C.Structure):
    _fields_ = [("version", C.c_float), ("per_coil", C.c_int), ("ncoils", C.c_int),
                ("nslices", C.c_int), ("total", C.c_int), ("nx", C.c_int),
                ("ny", C.c_int), ("nz", C.c_int)]


_lib = None


def load() -> C.CDLL:
    """Load libimpatient.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `make -C {PKG_ROOT}` "
                "(the HIP extension is required; there is no fallback)")
        lib = C.CDLL(str(LIB_PATH))
        lib.imp_last_error.restype = C.c_char_p
        lib.imp_version_string.restype = C.c_char_p
        lib.imp_status_string.restype = C.c_char_p
        _lib = lib
    return _lib


def last_error() -> str:
    return load().imp_last_error().decode(errors="replace")


def check(status: int, where: str) -> None:
    if status != IMP_OK:
        raise ImpError(status, where, last_error())


def call(name: str, *args) -> None:
    fn = getattr(load(), name)
    check(fn(*args), name)


def ptr(a) -> C.c_void_p:
    """Device pointer of a DeviceArray, or NULL for None."""
    if a is None:
        return C.c_void_p(0)
    return C.c_void_p(a.ptr)


class Context:
    """One HIP stream + workspace on one GPU (imp_ctx)."""

    def __init__(self, device: int = 0):
        lib = load()
        h = C.c_void_p()
        check(lib.imp_ctx_create(C.c_int(device), C.byref(h)), "imp_ctx_create")
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            st = getattr(self, "_staging", None)
            if st is not None:
                st.free()
                self._staging = None
            dst = getattr(self, "_dstaging", None)
            if dst is not None:
                dst.free()
                self._dstaging = None
            load().imp_ctx_destroy(self.h)
            self.h = None

    def staging(self, n: int) -> np.ndarray:
        """A page-locked float32 host array of at least n values, kept by
        the context and grown on demand (uploads from it run at the link's
        DMA rate).  Its contents are valid until the next call."""
        st = getattr(self, "_staging", None)
        if st is None or st.size < n:
            if st is not None:
                st.free()
            self._staging = st = HostStaging(self, max(n, 1))
        return st.array[:n]

    def device_staging(self, n: int) -> "DeviceArray":
        """A float32 device buffer of at least n values, kept by the context
        and grown on demand (the device side of staging(): no allocation or
        free, which synchronises the device, per upload).  Its contents are
        valid until the next call."""
        d = getattr(self, "_dstaging", None)
        if d is None or d.shape[0] < n:
            if d is not None:
                d.free()
            self._dstaging = d = DeviceArray(self, (max(n, 1),))
        return d

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        call("imp_ctx_synchronize", self.h)

    def stream(self) -> int:
        s = C.c_void_p()
        call("imp_ctx_stream", self.h, C.byref(s))
        return s.value or 0

    # -------------------------------------------------------------- memory
    def alloc(self, shape, dtype=np.float32) -> "DeviceArray":
        return DeviceArray(self, shape, dtype)

    def to_device(self, a) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.shape, a.dtype)
        d.upload(a)
        return d


class DeviceArray:
    """A device buffer owned through imp_device_alloc / imp_device_free."""

    def __init__(self, ctx: Context, shape, dtype=np.float32):
        self.ctx = ctx
        self.shape = tuple(np.atleast_1d(shape).tolist()) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        call("imp_device_alloc", ctx.h, C.c_size_t(self.nbytes), C.byref(p))
        self.ptr = p.value

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        call("imp_copy_h2d", self.ctx.h, C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p),
             C.c_size_t(self.nbytes))

    def get(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        call("imp_copy_d2h", self.ctx.h, out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
             C.c_size_t(self.nbytes))
        return out

    def zero(self):
        call("imp_memset_d", self.ctx.h, C.c_void_p(self.ptr), C.c_int(0), C.c_size_t(self.nbytes))

    def offset(self, elems: int) -> "DeviceView":
        return DeviceView(self, elems)

    def free(self):
        if self.ptr:
            load().imp_device_free(self.ctx.h, C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                self.free()
        except Exception:
            pass


class HostStaging:
    """Page-locked host buffer (imp_host_alloc / imp_host_free) of float32."""

    def __init__(self, ctx: Context, size: int):
        self.ctx = ctx
        self.size = int(size)
        p = C.c_void_p()
        call("imp_host_alloc", ctx.h, C.c_size_t(4 * self.size), C.byref(p))
        self.ptr = p.value
        self.array = np.ctypeslib.as_array((C.c_float * self.size).from_address(self.ptr))

    def free(self):
        if self.ptr:
            self.array = None
            load().imp_host_free(self.ctx.h, C.c_void_p(self.ptr))
            self.ptr = None


class DeviceView:
    """Pointer into a DeviceArray at an element offset (no ownership)."""

    def __init__(self, base: DeviceArray, elems: int):
        self.base = base
        self.ptr = base.ptr + elems * base.dtype.itemsize
