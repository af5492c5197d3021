"""NumPy-facing wrappers of the device operators (upload, call, download).

Every function here runs the HIP kernels of libimpatient.so through the C
ABI; none has a host fallback.  Complex arrays are numpy complex64/128 and
are split into the float32 real / imaginary arrays the C ABI takes.
"""
from __future__ import annotations

import ctypes as C
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from ._lib import (CGStats, Context, DeviceArray, IMP_CG_BEST_ITERATE, call, load, ptr)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _split(ctx: Context, z):
    z = np.asarray(z)
    return ctx.to_device(_f32(z.real)), ctx.to_device(_f32(z.imag))


def _opt(ctx: Context, a):
    return None if a is None else ctx.to_device(_f32(a))


def _pair(ctx: Context, n):
    return ctx.alloc((n,)), ctx.alloc((n,))


def _cplx(re: DeviceArray, im: DeviceArray, shape=None):
    z = re.get().astype(np.float32) + 1j * im.get().astype(np.float32)
    z = z.astype(np.complex64)
    return z.reshape(shape) if shape is not None else z


def nudft_adjoint(ctx: Context, kx, ky, kz, t, x, y, z, fm, d, sens=None, phi=None):
    """sum_c conj(S_c) sum_k conj(phi) d_c(k) e^{+i(2pi k.x + fm t)}; d is [C, M]."""
    d = np.atleast_2d(np.asarray(d))
    ncoils, nk = d.shape
    nx = len(x)
    a = [ctx.to_device(_f32(v)) for v in (kx, ky, t, x, y)]
    akz, az, afm = _opt(ctx, kz), _opt(ctx, z), _opt(ctx, fm)
    dr, di = _split(ctx, d.reshape(-1))
    sr = si = pr = pi = None
    if sens is not None:
        sr, si = _split(ctx, np.asarray(sens).reshape(-1))
    if phi is not None:
        pr, pi = _split(ctx, np.asarray(phi).reshape(-1))
    orr, oi = _pair(ctx, nx)
    call("imp_nudft_adjoint", ctx.h, C.c_int(nk), C.c_int(nx), C.c_int(ncoils),
         ptr(a[0]), ptr(a[1]), ptr(akz), ptr(a[2]), ptr(a[3]), ptr(a[4]), ptr(az), ptr(afm),
         ptr(pr), ptr(pi), ptr(dr), ptr(di), ptr(sr), ptr(si), ptr(orr), ptr(oi))
    return _cplx(orr, oi)


def _fill(host, parts, piece=1 << 22):
    """host[off:off + v.size] = v for every (off, v), in pieces of up to
    `piece` values on a thread pool (numpy releases the GIL while copying;
    the copies are strided reads of complex arrays plus first touches of
    host pages, several times faster spread over cores)."""
    tasks = [(off + i, v[i:i + piece]) for off, v in parts for i in range(0, v.size, piece)]

    def run(t):
        o, v = t
        host[o:o + v.size] = v

    workers = min(16, os.cpu_count() or 1, len(tasks))
    if workers <= 1:
        for t in tasks:
            run(t)
        return
    with ThreadPoolExecutor(workers) as ex:
        list(ex.map(run, tasks))


def nudft_adjoint_batch(ctx: Context, jobs, chunk_scale=0):
    """F^H d of several 2D images in one persistent launch.  jobs: list of
    dicts with kx, ky, t, x, y, fm, d [C, M], sens [C, N] (numpy).  Returns
    the list of images."""
    from ._lib import NudftJob
    # Every input goes into the context's page-locked staging buffer and its
    # device staging buffer (one upload at the DMA rate, no allocation per
    # array or call); fields start on 256-byte boundaries.
    def padded(n):
        return -(-n // 64) * 64

    layout = []
    n_in = n_out = 0
    for jb in jobs:
        d = np.atleast_2d(np.asarray(jb["d"]))
        ncoils, nk = d.shape
        nx = len(jb["x"])
        parts = [(k, _f32(jb[k])) for k in ("kx", "ky", "t", "x", "y", "fm")]
        dflat = d.reshape(-1)
        parts += [("d_r", dflat.real), ("d_i", dflat.imag)]
        if jb.get("sens") is not None:
            s = np.asarray(jb["sens"]).reshape(-1)
            parts += [("s_r", s.real), ("s_i", s.imag)]
        fields = {}
        for k, v in parts:
            fields[k] = (n_in, v)
            n_in += padded(v.size)
        layout.append((nk, nx, ncoils, fields, n_out))
        n_out += 2 * padded(nx)
    host = ctx.staging(max(n_in, 1))
    _fill(host, [f for _, _, _, fields, _ in layout for f in fields.values()])
    dev_in = ctx.device_staging(host.size)
    call("imp_copy_h2d", ctx.h, C.c_void_p(dev_in.ptr), host.ctypes.data_as(C.c_void_p),
         C.c_size_t(host.nbytes))
    dev_out = ctx.alloc((max(n_out, 1),))
    arr = (NudftJob * len(jobs))()
    for j, (nk, nx, ncoils, fields, o) in enumerate(layout):
        J = arr[j]
        J.num_k, J.num_x, J.ncoils = nk, nx, ncoils
        for k, (off, _) in fields.items():
            setattr(J, k, dev_in.ptr + 4 * off)
        J.out_r = dev_out.ptr + 4 * o
        J.out_i = dev_out.ptr + 4 * (o + padded(nx))
    call("imp_nudft_adjoint_batch", ctx.h, C.c_int(len(jobs)), arr, C.c_int(chunk_scale))
    res = dev_out.get()
    return [(res[o:o + nx] + 1j * res[o + padded(nx):o + padded(nx) + nx]).astype(np.complex64)
            for nk, nx, ncoils, fields, o in layout]


def nudft_forward(ctx: Context, kx, ky, kz, t, x, y, z, fm, rho, sens=None, ncoils=None):
    """d_c(k) = sum_x S_c(x) rho(x) e^{-i(2pi k.x + fm t)}; returns [C, M]."""
    nk, nx = len(kx), len(x)
    if ncoils is None:
        ncoils = 1 if sens is None else int(np.asarray(sens).size // nx)
    a = [ctx.to_device(_f32(v)) for v in (kx, ky, t, x, y)]
    akz, az, afm = _opt(ctx, kz), _opt(ctx, z), _opt(ctx, fm)
    rr, ri = _split(ctx, np.asarray(rho).reshape(-1))
    sr = si = None
    if sens is not None:
        sr, si = _split(ctx, np.asarray(sens).reshape(-1))
    dr, di = _pair(ctx, ncoils * nk)
    call("imp_nudft_forward", ctx.h, C.c_int(nk), C.c_int(nx), C.c_int(ncoils),
         ptr(a[0]), ptr(a[1]), ptr(akz), ptr(a[2]), ptr(a[3]), ptr(a[4]), ptr(az), ptr(afm),
         ptr(rr), ptr(ri), ptr(sr), ptr(si), ptr(dr), ptr(di))
    return _cplx(dr, di, (ncoils, nk))


def toeplitz_q_device(ctx: Context, kx, ky, t, Nx, Ny, ntime_segs):
    """Q_l of a 2D trajectory on the standard 2x grid, left on the device:
    returns the real and imaginary planes [(L+1) * 4 Nx Ny] as DeviceArrays
    (what imp_toep_create takes)."""
    a = [ctx.to_device(_f32(v)) for v in (kx, ky, t)]
    nq = 4 * Nx * Ny
    qr, qi = _pair(ctx, (ntime_segs + 1) * nq)
    call("imp_toeplitz_q", ctx.h, C.c_int(len(kx)), ptr(a[0]), ptr(a[1]), None, ptr(a[2]),
         None, None, C.c_int(Nx), C.c_int(Ny), C.c_int(1), C.c_int(nq), None, None, None,
         C.c_int(ntime_segs), C.c_int(0), ptr(qr), ptr(qi))
    return qr, qi


def toeplitz_q(ctx: Context, kx, ky, kz, t, Nx, Ny, Nz, ntime_segs, phi=None, flags=0,
               grid=None):
    """Q_l on the standard 2x grid (or on `grid` = (gx, gy, gz)); returns [(L+1), G]."""
    nk = len(kx)
    a = [ctx.to_device(_f32(v)) for v in (kx, ky, t)]
    akz = _opt(ctx, kz)
    pr = pi = None
    if phi is not None:
        pr, pi = _split(ctx, np.asarray(phi).reshape(-1))
    if grid is None:
        nq = 4 * Nx * Ny * (2 * Nz if Nz > 1 else 1)
        g = (None, None, None)
    else:
        nq = len(grid[0])
        g = tuple(_opt(ctx, v) for v in grid)
    nseg = ntime_segs + 1
    qr, qi = _pair(ctx, nseg * nq)
    call("imp_toeplitz_q", ctx.h, C.c_int(nk), ptr(a[0]), ptr(a[1]), ptr(akz), ptr(a[2]),
         ptr(pr), ptr(pi), C.c_int(Nx), C.c_int(Ny), C.c_int(Nz), C.c_int(nq),
         ptr(g[0]), ptr(g[1]), ptr(g[2]), C.c_int(ntime_segs), C.c_int(flags), ptr(qr), ptr(qi))
    return _cplx(qr, qi, (nseg, nq))


def toeplitz_q_gridding(ctx: Context, kx, ky, kz, t, Nx, Ny, Nz, ntime_segs, gridOS):
    """Q_l by Kaiser-Bessel gridding on the standard 2x grid; [(L+1), G]."""
    nk = len(kx)
    a = [ctx.to_device(_f32(v)) for v in (kx, ky, t)]
    akz = _opt(ctx, kz)
    nq = 4 * Nx * Ny * (2 * Nz if Nz > 1 else 1)
    nseg = ntime_segs + 1
    qr, qi = _pair(ctx, nseg * nq)
    call("imp_toeplitz_q_gridding", ctx.h, C.c_int(nk), ptr(a[0]), ptr(a[1]), ptr(akz),
         ptr(a[2]), C.c_int(Nx), C.c_int(Ny), C.c_int(Nz), C.c_int(ntime_segs),
         C.c_float(gridOS), ptr(qr), ptr(qi))
    return _cplx(qr, qi, (nseg, nq))


def fhd_gridding(ctx: Context, kx, ky, kz, t, Nx, Ny, Nz, fm, d, sens, ntime_segs, gridOS):
    """F^H d by gridding with time segmentation; d is [C, M]; returns [N]."""
    d = np.atleast_2d(np.asarray(d))
    ncoils, nk = d.shape
    a = [ctx.to_device(_f32(v)) for v in (kx, ky, t)]
    akz, afm = _opt(ctx, kz), _opt(ctx, fm)
    dr, di = _split(ctx, d.reshape(-1))
    sr = si = None
    if sens is not None:
        sr, si = _split(ctx, np.asarray(sens).reshape(-1))
    nv = Nx * Ny * Nz
    orr, oi = _pair(ctx, nv)
    call("imp_fhd_gridding", ctx.h, C.c_int(nk), C.c_int(ncoils), ptr(a[0]), ptr(a[1]),
         ptr(akz), ptr(a[2]), C.c_int(Nx), C.c_int(Ny), C.c_int(Nz), ptr(afm), ptr(dr),
         ptr(di), ptr(sr), ptr(si), C.c_int(ntime_segs), C.c_float(gridOS), ptr(orr), ptr(oi))
    return _cplx(orr, oi)


def fd_reg(ctx: Context, p, lam, w=None):
    """lambda * sum_d D_d^T W D_d p for p of shape (N1, N2) or (N1, N2, N3)."""
    p = np.asarray(p)
    shp = p.shape + (1,) * (3 - p.ndim)
    pr, pi = _split(ctx, p.reshape(-1))
    aw = _opt(ctx, None if w is None else np.asarray(w).reshape(-1))
    sr, si = _pair(ctx, p.size)
    call("imp_fd_reg_apply", ctx.h, C.c_int(shp[0]), C.c_int(shp[1]), C.c_int(shp[2]),
         C.c_float(lam), ptr(pr), ptr(pi), ptr(aw), ptr(sr), ptr(si))
    return _cplx(sr, si, p.shape)


def tv_weight(ctx: Context, x):
    x = np.asarray(x)
    shp = x.shape + (1,) * (3 - x.ndim)
    xr, xi = _split(ctx, x.reshape(-1))
    w = ctx.alloc((x.size,))
    call("imp_tv_weight", ctx.h, C.c_int(shp[0]), C.c_int(shp[1]), C.c_int(shp[2]),
         ptr(xr), ptr(xi), ptr(w))
    return w.get().reshape(x.shape)


class Toeplitz:
    """Resident Toeplitz normal operator (imp_toep)."""

    def __init__(self, ctx: Context, shape, q, tau, t0, fm=None, sens=None, lam=0.0):
        """q: the L+1 planes on the host ([L+1, grid] complex), or already on
        the device as a (real, imaginary) pair of DeviceArrays."""
        shape = tuple(shape) + (1,) * (3 - len(shape))
        self.ctx, self.shape = ctx, shape
        self.nvox = int(np.prod(shape))
        if isinstance(q, tuple):
            self._qr, self._qi = q
            grid = self.nvox * (8 if shape[2] > 1 else 4)
            nseg = int(np.prod(self._qr.shape)) // grid
        else:
            q = np.asarray(q)
            nseg = q.shape[0]
            self._qr, self._qi = _split(ctx, q.reshape(-1))
        self._fm = _opt(ctx, None if fm is None else np.asarray(fm).reshape(-1))
        ncoils = 1
        self._sr = self._si = None
        if sens is not None:
            sens = np.asarray(sens).reshape(-1, self.nvox)
            ncoils = sens.shape[0]
            self._sr, self._si = _split(ctx, sens.reshape(-1))
        h = C.c_void_p()
        call("imp_toep_create", ctx.h, C.c_int(shape[0]), C.c_int(shape[1]), C.c_int(shape[2]),
             C.c_int(ncoils), C.c_int(nseg - 1), C.c_float(tau), C.c_float(t0), ptr(self._qr),
             ptr(self._qi), ptr(self._fm), ptr(self._sr), ptr(self._si), C.c_float(lam),
             C.byref(h))
        self.h = h
        self._in = _pair(ctx, self.nvox)
        self._out = _pair(ctx, self.nvox)

    def close(self):
        if getattr(self, "h", None):
            call("imp_toep_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def apply(self, p):
        p = np.asarray(p).reshape(-1)
        self._in[0].upload(_f32(p.real))
        self._in[1].upload(_f32(p.imag))
        call("imp_toep_apply", self.h, ptr(self._in[0]), ptr(self._in[1]), ptr(self._out[0]),
             ptr(self._out[1]))
        return _cplx(*self._out)

    def cg(self, b, max_iter, tol=0.0, flags=IMP_CG_BEST_ITERATE, x0=None):
        b = np.asarray(b).reshape(-1)
        self._in[0].upload(_f32(b.real))
        self._in[1].upload(_f32(b.imag))
        if x0 is not None:
            x0 = np.asarray(x0).reshape(-1)
            self._out[0].upload(_f32(x0.real))
            self._out[1].upload(_f32(x0.imag))
        st = CGStats()
        call("imp_toep_cg", self.h, ptr(self._in[0]), ptr(self._in[1]), ptr(self._out[0]),
             ptr(self._out[1]), C.c_int(max_iter), C.c_double(tol), C.c_int(flags),
             C.byref(st))
        return _cplx(*self._out), st.as_dict()


    def set_sparse_reg(self, reg=None):
        """Use lambda C^T W C of a SparseReg as R (None: the FD stencil).
        The operator keeps a reference, so reg stays alive while in use."""
        call("imp_toep_set_sparse_reg", self.h, reg.h if reg is not None else None)
        self._reg = reg

    def precond_weights(self):
        """Current circulant preconditioner weights (imp_toep_get_precond),
        shaped like the image, natural DFT order."""
        w = self.ctx.alloc((self.nvox,))
        call("imp_toep_get_precond", self.h, ptr(w))
        return w.get().reshape(self.shape)

    def set_precond(self, w=None):
        """Set the circulant weights (None restores the Chan weights)."""
        if w is None:
            call("imp_toep_set_precond", self.h, None)
            return
        dw = self.ctx.to_device(_f32(np.asarray(w).reshape(-1)))
        call("imp_toep_set_precond", self.h, ptr(dw))

    def precond_apply(self, r):
        r = np.asarray(r).reshape(-1)
        self._in[0].upload(_f32(r.real))
        self._in[1].upload(_f32(r.imag))
        call("imp_toep_precond_apply", self.h, ptr(self._in[0]), ptr(self._in[1]),
             ptr(self._out[0]), ptr(self._out[1]))
        return _cplx(*self._out)


def read_mtx(path):
    """Matrix Market coordinate file (imp_read_mtx_size / imp_read_mtx, host
    only): (num_rows, num_cols, I, J, V) with 0-based int32 indices."""
    r, c, n = C.c_int(), C.c_int(), C.c_int()
    p = str(path).encode()
    call("imp_read_mtx_size", p, C.byref(r), C.byref(c), C.byref(n))
    I = np.zeros(max(n.value, 1), np.int32)
    J = np.zeros(max(n.value, 1), np.int32)
    V = np.zeros(max(n.value, 1), np.float32)
    call("imp_read_mtx", p, C.c_int(n.value), I.ctypes.data_as(C.c_void_p),
         J.ctypes.data_as(C.c_void_p), V.ctypes.data_as(C.c_void_p))
    return r.value, c.value, I[:n.value], J[:n.value], V[:n.value]


class SparseReg:
    """Sparse-matrix regularizer lambda C^T W C (imp_sparse_reg_*), C given
    as 0-based COO (rows, cols, I, J, V)."""

    def __init__(self, ctx: Context, num_rows, num_cols, I, J, V):
        I = np.ascontiguousarray(I, np.int32)
        J = np.ascontiguousarray(J, np.int32)
        V = np.ascontiguousarray(V, np.float32)
        # the library reads len(V) entries from each array
        if I.ndim != 1 or J.ndim != 1 or V.ndim != 1 or not (I.size == J.size == V.size):
            raise ValueError(f"SparseReg: I, J, V must be 1-D of one length "
                             f"(got {I.shape}, {J.shape}, {V.shape})")
        self.ctx = ctx
        self.rows, self.cols = int(num_rows), int(num_cols)
        h = C.c_void_p()
        call("imp_sparse_reg_create", ctx.h, C.c_int(self.rows), C.c_int(self.cols),
             C.c_int(len(V)), I.ctypes.data_as(C.c_void_p), J.ctypes.data_as(C.c_void_p),
             V.ctypes.data_as(C.c_void_p), C.byref(h))
        self.h = h

    def apply(self, p, lam, w=None):
        p = np.asarray(p)
        pr, pi = _split(self.ctx, p.reshape(-1))
        aw = _opt(self.ctx, None if w is None else np.asarray(w).reshape(-1))
        sr, si = _pair(self.ctx, self.cols)
        call("imp_sparse_reg_apply", self.h, C.c_float(lam), ptr(aw), ptr(pr), ptr(pi),
             ptr(sr), ptr(si))
        return _cplx(sr, si, p.shape)

    def close(self):
        if getattr(self, "h", None):
            call("imp_sparse_reg_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bruteforce_cg(ctx: Context, kx, ky, t, x, y, fm, d, sens, cg_num, lam, x0, reg=None,
                  tv_num=1):
    """bruteForceGpu's CG from x0 (cg_num - 1 iterations); d is [C, M].
    reg: a SparseReg for R = C^T W C (the -reg branch) instead of the stencil;
    tv_num > 1: total-variation passes (W from the current image)."""
    d = np.atleast_2d(np.asarray(d))
    ncoils, nk = d.shape
    nx = len(x)
    a = [ctx.to_device(_f32(v)) for v in (kx, ky, t, x, y)]
    afm = _opt(ctx, fm)
    dr, di = _split(ctx, d.reshape(-1))
    sr = si = None
    if sens is not None:
        sr, si = _split(ctx, np.asarray(sens).reshape(-1))
    xr, xi = _split(ctx, np.asarray(x0).reshape(-1))
    head = (ctx.h, C.c_int(nk), C.c_int(nx), C.c_int(ncoils), ptr(a[0]), ptr(a[1]),
            ptr(a[2]), ptr(a[3]), ptr(a[4]), ptr(afm), ptr(dr), ptr(di), ptr(sr), ptr(si),
            C.c_int(cg_num), C.c_float(lam))
    if reg is None and tv_num <= 1:
        call("imp_bruteforce_cg", *head, ptr(xr), ptr(xi))
    else:
        call("imp_bruteforce_cg_ex", *head, reg.h if reg is not None else None,
             C.c_int(tv_num), ptr(xr), ptr(xi))
    return _cplx(xr, xi)


class CNN:
    """Learned warm-start network (imp_cnn_*): x = z + s * net(z / s)."""

    def __init__(self, ctx: Context, H: int, W: int, seed: int = 1):
        self.ctx, self.H, self.W = ctx, H, W
        self.n = H * W
        h = C.c_void_p()
        call("imp_cnn_create", ctx.h, C.c_int(H), C.c_int(W), C.c_ulonglong(seed), C.byref(h))
        self.h = h
        lib = load()
        lib.imp_cnn_num_params.restype = C.c_int
        self.nparam = int(lib.imp_cnn_num_params())
        self._z = _pair(ctx, self.n)
        self._t = _pair(ctx, self.n)
        self._x = _pair(ctx, self.n)

    def close(self):
        if getattr(self, "h", None):
            call("imp_cnn_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        out = np.zeros(self.nparam, np.float32)
        call("imp_cnn_get_params", self.h, out.ctypes.data_as(C.c_void_p), C.c_int(self.nparam))
        return out

    def set_params(self, params):
        params = _f32(np.asarray(params).reshape(-1))
        call("imp_cnn_set_params", self.h, params.ctypes.data_as(C.c_void_p),
             C.c_int(self.nparam))

    def _load(self, pair, v):
        v = np.asarray(v).reshape(-1)
        pair[0].upload(_f32(v.real))
        pair[1].upload(_f32(v.imag))

    def forward(self, z):
        self._load(self._z, z)
        call("imp_cnn_forward", self.h, ptr(self._z[0]), ptr(self._z[1]), ptr(self._x[0]),
             ptr(self._x[1]))
        return _cplx(*self._x).reshape(self.H, self.W)

    def gradient(self, z, t):
        self._load(self._z, z)
        self._load(self._t, t)
        g = np.zeros(self.nparam, np.float32)
        loss = C.c_double()
        call("imp_cnn_gradient", self.h, ptr(self._z[0]), ptr(self._z[1]), ptr(self._t[0]),
             ptr(self._t[1]), g.ctypes.data_as(C.c_void_p), C.byref(loss))
        return loss.value, g

    def train_step_device(self, zr, zi, tr, ti, lr):
        """One Adam step on device-resident (z, t) (DeviceArray re/im pairs)."""
        loss = C.c_double()
        call("imp_cnn_train_step", self.h, ptr(zr), ptr(zi), ptr(tr), ptr(ti), C.c_float(lr),
             C.byref(loss))
        return loss.value

    def forward_device(self, zr, zi, xr, xi):
        call("imp_cnn_forward", self.h, ptr(zr), ptr(zi), ptr(xr), ptr(xi))

