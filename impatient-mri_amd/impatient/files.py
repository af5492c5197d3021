"""Reference data files through the C ABI readers / writers (host only)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from ._lib import FileInfo, call, load


def info(path) -> FileInfo:
    fi = FileInfo()
    call("imp_read_data_file", str(path).encode(), C.byref(fi), None, C.c_int(0))
    return fi


def read(path, with_info: bool = False):
    """Read a v0.2 or v1.0 data file; returns float32 array (and FileInfo)."""
    fi = info(path)
    buf = np.empty(max(fi.total, 1), dtype=np.float32)
    call("imp_read_data_file", str(path).encode(), C.byref(fi),
         buf.ctypes.data_as(C.POINTER(C.c_float)), C.c_int(buf.size))
    out = buf[:fi.total]
    return (out, fi) if with_info else out


def write_v02(path, data, per_coil=None, ncoils=1, nslices=1):
    a = np.ascontiguousarray(data, dtype=np.float32).reshape(-1)
    per = a.size // ncoils if per_coil is None else per_coil
    call("imp_write_data_file_v02", str(path).encode(), C.c_int(per), C.c_int(ncoils),
         C.c_int(nslices), a.ctypes.data_as(C.POINTER(C.c_float)), C.c_int(a.size))


def write_v10(path, data, nx, ny, nz, ncoils=-1, nslices=1):
    a = np.ascontiguousarray(data, dtype=np.float32).reshape(-1)
    call("imp_write_data_file_v10", str(path).encode(), C.c_int(nx), C.c_int(ny), C.c_int(nz),
         C.c_int(ncoils), C.c_int(nslices), a.ctypes.data_as(C.POINTER(C.c_float)),
         C.c_int(a.size))


def write_problem(dirpath, prob, kdata, version: float = 1.0):
    """Write a synth.Problem and its k-space data [C, M] as an input directory
    with the reference file names (coil_number always present for per-coil data)."""
    d = Path(dirpath)
    d.mkdir(parents=True, exist_ok=True)
    nk, nv, C_ = prob.nk, prob.nvox, prob.ncoils
    z = lambda n: np.zeros(n, np.float32)
    fields = {
        "kx": (prob.kx, -1), "ky": (prob.ky, -1),
        "kz": (prob.kz if prob.kz is not None else z(nk), -1), "t": (prob.t, -1),
        "ix": (prob.x, -1), "iy": (prob.y, -1),
        "iz": (prob.z if prob.z is not None else z(nv), -1), "fm": (prob.fm, -1),
        "kdata_r": (np.asarray(kdata).real.reshape(-1), C_),
        "kdata_i": (np.asarray(kdata).imag.reshape(-1), C_),
        "idata_r": (z(nv), -1), "idata_i": (z(nv), -1),
    }
    if prob.sens is not None:
        fields["sensi_r"] = (prob.sens.real.reshape(-1), C_)
        fields["sensi_i"] = (prob.sens.imag.reshape(-1), C_)
    else:
        fields["sensi_r"] = (np.ones(nv, np.float32), 1)
        fields["sensi_i"] = (z(nv), 1)
    for name, (arr, nc) in fields.items():
        p = d / f"{name}.dat"
        if version >= 1.0:
            write_v10(p, arr, prob.Nx, prob.Ny, prob.Nz, nc)
        else:
            write_v02(p, arr, ncoils=max(nc, 1))


__all__ = ["info", "read", "write_v02", "write_v10", "write_problem", "load"]
