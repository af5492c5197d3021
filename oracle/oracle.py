"""FP64 CPU oracle of the IMPATIENT hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker.  The product (libimpatient.so) never
calls it.

Each function restates the reference algorithm it cites (paths are relative
to the reference repository) in float64 numpy.  How each is pinned (closed
forms, identities, independent restatements) is listed in
tests/test_oracle.py; every function below has at least one such check.

Conventions (README "Data Format"; kernels/ftGpu.cu:493-498):
  forward  d_c(k) = sum_x S_c(x) rho(x) exp(-i(2 pi k.x + w(x) t_k))
  adjoint  rho(x) = sum_c conj(S_c(x)) sum_k d_c(k) exp(+i(2 pi k.x + w(x) t_k))
"""
from __future__ import annotations

import numpy as np

_BLOCK = 1 << 22   # complex elements per phase block (memory bound of the oracle)


def _phase(kx, ky, kz, t, x, y, z, fm, k_sl):
    """2 pi k.x + w t for samples k_sl (rows) and all voxels (cols), FP64."""
    ph = 2.0 * np.pi * (np.outer(kx[k_sl], x) + np.outer(ky[k_sl], y))
    if kz is not None and z is not None:
        ph += 2.0 * np.pi * np.outer(kz[k_sl], z)
    if fm is not None:
        ph += np.outer(t[k_sl], fm)
    return ph


def _as64(*arrs):
    return [None if a is None else np.asarray(a, dtype=np.float64) for a in arrs]


def nudft_forward(kx, ky, kz, t, x, y, z, fm, rho, sens=None):
    """ftCpu (kernels/ftCpu.cpp:110-189) with the SENSE wrap of parallelFtCpu
    (kernels/parImagingCpu.cpp:85-112): d_c(k) = sum_x S_c(x) rho(x) e^{-i phase}.
    Returns [C, M] complex128."""
    kx, ky, kz, t, x, y, z, fm = _as64(kx, ky, kz, t, x, y, z, fm)
    rho = np.asarray(rho, dtype=np.complex128).reshape(-1)
    S = np.ones((1, len(x))) if sens is None else np.asarray(sens, np.complex128).reshape(-1, len(x))
    src = S * rho[None, :]                                   # [C, N]
    nk = len(kx)
    out = np.zeros((S.shape[0], nk), dtype=np.complex128)
    step = max(1, _BLOCK // max(1, len(x)))
    for k0 in range(0, nk, step):
        sl = slice(k0, min(nk, k0 + step))
        E = np.exp(-1j * _phase(kx, ky, kz, t, x, y, z, fm, sl))   # [m, N]
        out[:, sl] = src @ E.T
    return out


def nudft_adjoint(kx, ky, kz, t, x, y, z, fm, d, sens=None, phi=None):
    """iftCpu (kernels/ftCpu.cpp:199-277) + parallelIftCpu
    (kernels/parImagingCpu.cpp:114-154) and the direct F^H d of
    ComputeFH_GPU_BF / parcomputeAH_GPU_BF (toeplitz/computeFH.cmem.cu:155-200,
    1146-1186), with the conj(phi) d weighting of ComputeRhoPhiGPU (:112-127).
    d is [C, M]; returns [N] complex128."""
    kx, ky, kz, t, x, y, z, fm = _as64(kx, ky, kz, t, x, y, z, fm)
    d = np.atleast_2d(np.asarray(d, dtype=np.complex128))
    if phi is not None:
        d = np.conj(np.atleast_2d(np.asarray(phi, np.complex128)).reshape(d.shape)) * d
    C = d.shape[0]
    S = np.ones((C, len(x))) if sens is None else np.asarray(sens, np.complex128).reshape(C, len(x))
    nk = len(kx)
    acc = np.zeros((C, len(x)), dtype=np.complex128)
    step = max(1, _BLOCK // max(1, len(x)))
    for k0 in range(0, nk, step):
        sl = slice(k0, min(nk, k0 + step))
        E = np.exp(1j * _phase(kx, ky, kz, t, x, y, z, fm, sl))    # [m, N]
        acc += d[:, sl] @ E
    return np.sum(np.conj(S) * acc, axis=0)


def hanning(tm, tau, l, t0):
    """hanning_d (toeplitz/computeQ.cmem.cu:125-138):
    1/2 + 1/2 cos(pi (t - tau l - t0)/tau) inside |t - tau l - t0| < tau, else 0."""
    a = np.asarray(tm, np.float64) - tau * l - t0
    return np.where(np.abs(a) < tau, 0.5 + 0.5 * np.cos(np.pi * a / tau), 0.0)


def segment_timing(t, ntime_segs):
    """tau, t0 as recon.cu:1785-1787 computes them (float32 arithmetic)."""
    t = np.asarray(t, np.float32)
    span = np.float32(t[-1] - t[0])
    tau = np.float32(span / np.float32(ntime_segs)) if ntime_segs > 0 else span
    return float(tau), float(t[0])


def q_grid(Nx, Ny, Nz):
    """Q-grid coordinates written by main_mri.cu:461-500: y outer, x, z
    inner; r = (x/Nx, y/Ny, z/Nz) with x in [-Nx, Nx) etc. (z absent in 2D)."""
    ys = np.arange(-Ny, Ny) / Ny
    xs = np.arange(-Nx, Nx) / Nx
    if Nz > 1:
        zs = np.arange(-Nz, Nz) / Nz
        Y, X, Z = np.meshgrid(ys, xs, zs, indexing="ij")
        return X.reshape(-1), Y.reshape(-1), Z.reshape(-1)
    Y, X = np.meshgrid(ys, xs, indexing="ij")
    return X.reshape(-1), Y.reshape(-1), np.zeros(X.size)


def toeplitz_q(kx, ky, kz, t, Nx, Ny, Nz, ntime_segs, phi=None, cumulative=False, grid=None):
    """Q_l(r) = sum_k a_l(t_k) |phi_k|^2 e^{+i 2 pi k.r}, l = 0..L
    (ComputeQ_GPU / toeplitz_computeQ_GPU, toeplitz/computeQ.cmem.cu:169-208,
    826-1028).  cumulative=True restates the reference direct path, whose
    accumulator is not cleared between segments (:190-191, 951-963).
    Returns [(L+1), G] complex128."""
    kx, ky, kz, t = _as64(kx, ky, kz, t)
    gx, gy, gz = q_grid(Nx, Ny, Nz) if grid is None else _as64(*grid)
    tau, t0 = segment_timing(t, ntime_segs)
    mag = np.ones(len(kx)) if phi is None else np.abs(np.asarray(phi, np.complex128)) ** 2
    W = np.stack([hanning(t, tau, l, t0) * mag for l in range(ntime_segs + 1)])  # [L+1, M]
    use_z = kz is not None and Nz > 1
    out = np.zeros((ntime_segs + 1, len(gx)), dtype=np.complex128)
    step = max(1, _BLOCK // max(1, len(gx)))
    for k0 in range(0, len(kx), step):
        sl = slice(k0, min(len(kx), k0 + step))
        ph = 2.0 * np.pi * (np.outer(kx[sl], gx) + np.outer(ky[sl], gy))
        if use_z:
            ph += 2.0 * np.pi * np.outer(kz[sl], gz)
        out += W[:, sl] @ np.exp(1j * ph)
    if cumulative:
        out = np.cumsum(out, axis=0)
    return out


def _shift(a, inverse):
    """cuda_fftshift / cuda_fft3shift on even grids (recon.cu:235-303,
    fftshift.cu:5-104): a swap of half-spaces along every axis."""
    return np.fft.ifftshift(a) if inverse else np.fft.fftshift(a)


def toeplitz_apply(p, q, tau, t0, fm=None, sens=None, lam=0.0, w=None):
    """q = sum_c conj(S_c) sum_l e^{+i w tau_l} T_l(e^{-i w tau_l} S_c p) + lam D^H W D p
    with T_l v = crop(IFFT(FFT(ifftshift(pad_center(v))) Qf_l)) / G and
    Qf_l = FFT(ifftshift(Q_l)): Computeq / parCudaAhA / CudaAhA / CudaFhF
    (toeplitz/recon.cu:1113-1202, 1370-1465, 1467-1697, Qf at :1939-1964),
    with the reference's centre padding (cuda_oversample :319-367) and
    centre crop (cuda_undersample_scale :381-421).
    p: image [N1, N2] or [N1, N2, N3]; q: [(L+1), G] on the standard grid."""
    p = np.asarray(p, np.complex128)
    shp = p.shape
    d3 = p.ndim == 3 and shp[2] > 1
    N = shp if d3 else shp[:2]
    M = tuple(2 * n for n in N)
    G = int(np.prod(M))
    q = np.asarray(q, np.complex128).reshape(q.shape[0], *M)
    Qf = [np.fft.fftn(_shift(q[l], True)) for l in range(q.shape[0])]
    f = np.zeros(N) if fm is None else np.asarray(fm, np.float64).reshape(N)
    S = [np.ones(N)] if sens is None else list(np.asarray(sens, np.complex128).reshape(-1, *N))
    mid = tuple(n - n // 2 for n in N)
    sl = tuple(slice(m, m + n) for m, n in zip(mid, N))
    pv = p.reshape(N)
    out = np.zeros(N, dtype=np.complex128)
    for Sc in S:
        acc = np.zeros(N, dtype=np.complex128)
        for l in range(len(Qf)):
            tl = tau * l + t0
            v = np.exp(-1j * f * tl) * (Sc * pv)
            big = np.zeros(M, dtype=np.complex128)
            big[sl] = v
            # unnormalised inverse FFT scaled by 1/G (cuda_undersample_scale)
            r = np.fft.ifftn(np.fft.fftn(_shift(big, True)) * Qf[l])
            r = _shift(r, False)[sl]
            acc += np.exp(1j * f * tl) * r
        out += np.conj(Sc) * acc
    if lam:
        out += fd_reg(pv, lam, w)
    return out.reshape(shp)


def toeplitz_operator(q, tau, t0, shape, fm=None, sens=None, lam=0.0, w=None, workers=-1):
    """toeplitz_apply as a callable v -> A v for iterated FP64 solves at
    configuration sizes: Qf_l computed once, all coils in one batched
    transform through scipy.fft (the same pocketfft transforms as numpy.fft,
    threaded), and the low-corner embedding.  The centre padding with
    ifftshift places v at offset -N/2 (mod 2N) and the crop after fftshift
    reads at the same offset; a circular convolution commutes with shifts,
    so placing v at 0 and reading at 0 is the same map.
    v and the result are flat [N] complex128."""
    import scipy.fft as sfft
    N = tuple(shape) if len(shape) == 3 and shape[2] > 1 else tuple(shape[:2])
    M = tuple(2 * n for n in N)
    axes = tuple(range(1, len(N) + 1))
    q = np.asarray(q, np.complex128).reshape(q.shape[0], *M)
    Qf = [sfft.fftn(_shift(q[l], True), workers=workers) for l in range(q.shape[0])]
    f = np.zeros(N) if fm is None else np.asarray(fm, np.float64).reshape(N)
    S = (np.ones((1,) + N, np.complex128) if sens is None
         else np.asarray(sens, np.complex128).reshape((-1,) + N))
    ph = [np.exp(1j * f * (tau * l + t0)) for l in range(len(Qf))]
    corner = (slice(None),) + tuple(slice(0, n) for n in N)
    big = np.zeros((S.shape[0],) + M, dtype=np.complex128)

    def apply(v):
        pv = np.asarray(v, np.complex128).reshape(N)
        sp = S * pv
        acc = np.zeros_like(sp)
        for l in range(len(Qf)):
            big[corner] = np.conj(ph[l]) * sp   # the rest of big stays zero
            X = sfft.fftn(big, axes=axes, workers=workers)
            X *= Qf[l]
            r = sfft.ifftn(X, axes=axes, workers=workers, overwrite_x=True)
            acc += ph[l] * r[corner]
        out = (np.conj(S) * acc).sum(axis=0)
        if lam:
            out += fd_reg(pv, lam, w)
        return out.reshape(-1)
    return apply


def fd_reg(p, lam, w=None):
    """s = lam * sum_d D_d^T W D_d p, periodic: (D_d p)(i) = p(i) - p(i+e_d),
    (D_d^T t)(i) = t(i) - t(i-e_d).  DHWD2dGpu / DHWD3dGpu
    (toeplitz/DHWD2dGpu_toeplitz.cu:132-387, toeplitz/DHWD3dGpu_toeplitz.cu:170-580);
    the same periodic D as matlab_utils/create_sparse_matrix/createDWithPeriodicBoundary.m."""
    p = np.asarray(p, np.complex128)
    W = np.ones(p.shape) if w is None else np.asarray(w, np.float64).reshape(p.shape)
    s = np.zeros_like(p)
    for ax in range(p.ndim):
        if p.shape[ax] == 1:
            continue
        t1 = W * (p - np.roll(p, -1, axis=ax))
        s += t1 - np.roll(t1, 1, axis=ax)
    return lam * s


def sparse_reg(p, lam, num_rows, num_cols, I, J, V, w=None):
    """s = lam * C^T W C p for a real sparse C (num_rows x num_cols, 0-based
    COO; duplicate entries add, as coo2Csr keeps them, structures.cpp:380-392).
    The -reg branch of Computeq (toeplitz/recon.cu:1543-1640) and of the
    brute-force CG (bruteForce/bruteForceGpu.cu:513-541, 579-609):
    Cf = C p (smvmGpu), each block of num_cols rows of Cf times W
    (pointMultGpu with the imaginary part zero), C^T Cf (smvmGpu on the
    transposed CSR), times lam (addGpu with alpha = lambda2)."""
    import scipy.sparse as sp
    p = np.asarray(p, np.complex128)
    Cm = sp.coo_matrix((np.asarray(V, np.float64), (np.asarray(I), np.asarray(J))),
                       shape=(num_rows, num_cols)).tocsr()
    cf = Cm @ p.reshape(-1)
    if w is not None:
        wv = np.asarray(w, np.float64).reshape(-1)
        cf = cf * wv[np.arange(num_rows) % num_cols]
    return (lam * (Cm.T @ cf)).reshape(p.shape)


def tv_weight(x):
    """w = 0.5 / sqrt(sum_d |D_d x|^2 + 1e-6): multiply2dGpu / multiply3dGpu
    (toeplitz/multiplyGpu_toeplitz.cu:134-172) on Dhori / Dverti / Dzaxis."""
    x = np.asarray(x, np.complex128)
    s = np.zeros(x.shape)
    for ax in range(x.ndim):
        if x.shape[ax] == 1:
            continue
        s += np.abs(x - np.roll(x, -1, axis=ax)) ** 2
    return 0.5 / np.sqrt(s + 1e-6)


def _vec_ops(vec_dtype):
    """Storage rounding of CG vectors: complex128 (exact oracle) or complex64,
    the FP32 vectors of the reference (FLOAT_T = float, kernels/tools.h:118-120;
    cufftComplex matrices in toeplitz/recon.cu) with FP64 dot products."""
    vd = np.dtype(vec_dtype)
    sd = np.float32 if vd == np.complex64 else np.float64
    st = lambda v: np.asarray(v).astype(vd)
    dot = lambda a, b: np.vdot(np.asarray(a, np.complex128), np.asarray(b, np.complex128)).real
    return st, dot, sd


def toeplitz_cg(apply, b, max_iter, tol=1e-16, best_iterate=True, vec_dtype=np.complex128):
    """The CG loop of toeplitz_recon (toeplitz/recon.cu:2011-2133):
    x0 = 0, r = b, p = r; alpha = beta/Re(p^H q); x += alpha p;
    r -= alpha q; beta2 = |r|^2; p = r + (beta2/beta) p; keep the iterate
    with the smallest |r| (reconBest); stop when sqrt(beta2/|b|^2) < tol.
    vec_dtype=complex64 stores the vectors in FP32 (scalars stay FP64).
    Returns (x, iterations, residual history)."""
    st, dot, sd = _vec_ops(vec_dtype)
    b = st(b)
    x = st(np.zeros_like(b))
    r = b.copy()
    p = r.copy()
    beta2 = dot(r, r)
    normb = beta2
    best, xbest = beta2, x.copy()
    hist = []
    it = 0
    while it < max_iter:
        q = st(apply(p))
        beta = beta2
        alpha = sd(beta / dot(p, q))
        x = st(x + alpha * p)
        r = st(r - alpha * q)
        beta2 = dot(r, r)
        p = st(r + sd(beta2 / beta) * p)
        hist.append(np.sqrt(beta2 / normb))
        it += 1
        if beta2 < best:
            best, xbest = beta2, x.copy()
        if np.sqrt(beta2 / normb) < tol:
            break
    return (xbest if best_iterate else x), it, hist


GRID_KERNEL_WIDTH = 4.0   # kernelWidth (computeQ.cmem.cu:455, computeFH.cmem.cu:495)


def gridding_params(img, gridOS):
    """Grid set-up of computeQ_GPU_Grid / computeFH_GPU_Grid
    (toeplitz/computeQ.cmem.cu:433-498, computeFH.cmem.cu:495-543) for a
    transform of image size img = (nx, ny, nz) (nz = 1 in 2D): gridOS is
    rounded in float32 so the grid is even in x/y (2D) or a multiple of 4 in
    z (3D); beta is Beatty's (Eqn. 5; the GPU path, :456-463); grid sizes
    are ceil(gridOS n), x and y bumped to even.  Returns (os, beta, grid)."""
    f = np.float32
    nx, ny, nz = (int(v) for v in img)
    os_ = f(gridOS)
    if nz == 1:
        os_ = f(f(2.0) * f(np.ceil(f(os_ * f(nx)) / f(2.0))) / f(nx))
    else:
        os_ = f(f(4.0) * f(np.ceil(f(os_ * f(nz)) / f(4.0))) / f(nz))
    W = GRID_KERNEL_WIDTH
    o = float(os_)
    beta = np.pi * np.sqrt((o - 0.5) ** 2 * (W * W * 4.0) / (o * o) - 0.8)
    g = [int(np.ceil(f(os_ * f(nx)))), int(np.ceil(f(os_ * f(ny))))]
    g = [v + (v % 2) for v in g]
    g.append(1 if nz == 1 else int(np.ceil(f(os_ * f(nz)))))
    return o, float(beta), tuple(g)


def kb_kernel(dist, beta, width=GRID_KERNEL_WIDTH):
    """Kaiser-Bessel weight I0(beta sqrt(1 - (2 d / W)^2)) / W for d <= W/2,
    else 0 (gridding_GPU_2D / gridding_Gold_2D, GPU_kernels.cu:290-313,
    CPU_kernels.cpp:160-171); d in units of the unoversampled grid."""
    d = np.asarray(dist, np.float64)
    arg = 1.0 - (2.0 * d / width) ** 2
    return np.where(d <= width / 2, np.i0(beta * np.sqrt(np.clip(arg, 0.0, None))) / width, 0.0)


def kb_deapod(u, beta, width=GRID_KERNEL_WIDTH):
    """Fourier transform of kb_kernel at image position u (cycles per grid
    sample, u in [-1/2, 1/2)): sin(sqrt(pi^2 W^2 u^2 - beta^2)) / sqrt(...),
    sinh(...) below zero (Deapodization2d_kernel, gridding_utils.cu:60-110)."""
    e = (np.pi * width * np.asarray(u, np.float64)) ** 2 - beta * beta
    out = np.empty_like(e)
    pos = e >= 0
    out[pos] = np.sinc(np.sqrt(e[pos]) / np.pi)           # sin(s)/s
    s = np.sqrt(-e[~pos])
    out[~pos] = np.sinh(s) / s
    return out


def gridding_adjoint(k, data, img, gridOS, beta=None):
    """Gridding estimate of sum_j data_j e^{+i 2 pi k_j . x} on the image
    grid x = (i - n/2)/n per axis, as computeFH_GPU_Grid and
    computeQ_GPU_Grid evaluate it (toeplitz/computeFH.cmem.cu:470-802,
    computeQ.cmem.cu:401-754; CUDA_interface.cu, GPU_kernels.cu:190-581,
    gridding_utils.cu):
      1. convolve the samples with the separable Kaiser-Bessel kernel onto
         the oversampled grid, sample j at s = os (k_j + n/2) per axis,
         grid points g in [0, G) within W/2 (unoversampled units), no wrap;
      2. ifftshift, unnormalised inverse FFT, fftshift, crop the centre n;
      3. divide by the kernel's transform per axis and by os per axis.
    k: [M, d] (d = 2 or 3) in cycles of the transform's image; data [M]
    complex (already time-weighted); img = (nx, ny, nz).  beta overrides
    Beatty's value (gridding_Gold_2D/3D use 18.5547, CPU_kernels.cpp:110).
    Returns [ny, nx] or [ny, nx, nz] complex128 (y outer, x, z inner)."""
    os_, beta_b, G = gridding_params(img, gridOS)
    beta = beta_b if beta is None else beta
    W = GRID_KERNEL_WIDTH
    d3 = img[2] > 1
    dims = 3 if d3 else 2
    n = [int(img[0]), int(img[1]), int(img[2])]
    grid = gridding_grid(k, data, img, gridOS, beta)
    # ifftshift, unnormalised inverse DFT, fftshift, centre crop
    X = np.fft.ifftn(np.fft.ifftshift(grid)) * grid.size
    X = np.fft.fftshift(X)
    sl = tuple(slice(G[a] // 2 - n[a] // 2, G[a] // 2 - n[a] // 2 + n[a]) for a in range(dims))
    X = X[sl]
    # deapodization, separable, and 1 / os per axis
    for a in range(dims):
        u = (np.arange(n[a]) - n[a] // 2) / n[a]
        c = kb_deapod(u, beta, W) * os_
        shape = [1] * dims
        shape[a] = n[a]
        X = X / c.reshape(shape)
    return np.swapaxes(X, 0, 1)


def gridding_grid(k, data, img, gridOS, beta=None):
    """Step 1 of gridding_adjoint alone: the oversampled grid, axes (x, y[, z])
    (gridding_Gold_2D / gridding_Gold_3D before their reorder,
    CPU_kernels.cpp:89-330; gridding_GPU_2D/3D, GPU_kernels.cu:190-581)."""
    os_, beta_b, G = gridding_params(img, gridOS)
    beta = beta_b if beta is None else beta
    W = GRID_KERNEL_WIDTH
    k = np.asarray(k, np.float64)
    data = np.asarray(data, np.complex128).reshape(-1)
    d3 = img[2] > 1
    dims = 3 if d3 else 2
    n = [int(img[0]), int(img[1]), int(img[2])]
    grid = np.zeros(G[:dims], np.complex128)
    half = W * os_ / 2.0
    r = int(np.ceil(half)) + 1
    offs = np.arange(-r, r + 1)
    idx, wts = [], []
    for a in range(dims):
        s = os_ * (k[:, a] + n[a] / 2.0)
        g = np.floor(s)[:, None] + offs[None, :]
        dist = np.abs(s[:, None] - g) / os_
        ok = (g >= 0) & (g < G[a]) & (dist <= W / 2)
        idx.append(np.where(ok, g, 0).astype(np.int64))
        wts.append(np.where(ok, kb_kernel(dist, beta, W), 0.0))
    if dims == 2:
        w = wts[0][:, :, None] * wts[1][:, None, :]
        ix = np.broadcast_to(idx[0][:, :, None], w.shape)
        iy = np.broadcast_to(idx[1][:, None, :], w.shape)
        np.add.at(grid, (ix.reshape(-1), iy.reshape(-1)), (w * data[:, None, None]).reshape(-1))
    else:
        w = wts[0][:, :, None, None] * wts[1][:, None, :, None] * wts[2][:, None, None, :]
        ix = np.broadcast_to(idx[0][:, :, None, None], w.shape)
        iy = np.broadcast_to(idx[1][:, None, :, None], w.shape)
        iz = np.broadcast_to(idx[2][:, None, None, :], w.shape)
        np.add.at(grid, (ix.reshape(-1), iy.reshape(-1), iz.reshape(-1)),
                  (w * data[:, None, None, None]).reshape(-1))
    return grid


def gridding_q(kx, ky, kz, t, Nx, Ny, Nz, ntime_segs, gridOS, beta=None):
    """Q_l by gridding (computeQ_GPU_Grid, toeplitz/computeQ.cmem.cu:401-754,
    called per l from toeplitz_computeQ_GPU :826-1028): the samples, with
    weight a_l(t) (and phi = 1: unit data, :518-525), at 2k on the 2N grid.
    Returns [(L+1), G] complex128 on the standard Q grid (toeplitz_q's)."""
    kx, ky, kz, t = _as64(kx, ky, kz, t)
    d3 = Nz > 1
    k = np.stack([2 * kx, 2 * ky] + ([2 * kz] if d3 else []), axis=1)
    img = (2 * Nx, 2 * Ny, 2 * Nz if d3 else 1)
    tau, t0 = segment_timing(t, ntime_segs)
    out = [gridding_adjoint(k, hanning(t, tau, l, t0), img, gridOS, beta).reshape(-1)
           for l in range(ntime_segs + 1)]
    return np.stack(out)


def gridding_fhd(kx, ky, kz, t, fm, d, sens, Nx, Ny, Nz, ntime_segs, gridOS):
    """F^H d by gridding with time segmentation (parcomputeAH_GPU_Grid /
    computeAH_GPU_Grid, toeplitz/computeFH.cmem.cu:1032-1117):
    sum_c conj(S_c) sum_l e^{+i w (t0 + l tau)} grid(a_l(t) d_c), the image on
    x = (i - N/2)/N.  d: [C, M]; sens [C, N] or None.  Returns [N] complex128
    in the [Ny][Nx][Nz] order."""
    kx, ky, kz, t = _as64(kx, ky, kz, t)
    d = np.atleast_2d(np.asarray(d, np.complex128))
    d3 = Nz > 1
    k = np.stack([kx, ky] + ([kz] if d3 else []), axis=1)
    img = (Nx, Ny, Nz)
    tau, t0 = segment_timing(t, ntime_segs)
    nv = Nx * Ny * Nz
    f = np.zeros(nv) if fm is None else np.asarray(fm, np.float64).reshape(-1)
    S = np.ones((d.shape[0], nv)) if sens is None else np.asarray(sens, np.complex128).reshape(-1, nv)
    out = np.zeros(nv, np.complex128)
    for c in range(d.shape[0]):
        acc = np.zeros(nv, np.complex128)
        for l in range(ntime_segs + 1):
            g = gridding_adjoint(k, hanning(t, tau, l, t0) * d[c], img, gridOS).reshape(-1)
            acc += np.exp(1j * f * (tau * l + t0)) * g
        out += np.conj(S[c]) * acc
    return out


def toeplitz_recon(make_apply, b, shape, num_iter, enable_tv=False, tv_num=1, restarts=1,
                   tol=1e-16, vec_dtype=np.complex128):
    """The restart / TV / CG loops of toeplitz_recon (toeplitz/recon.cu:2011-2133).
    make_apply(w) returns v -> A v with the TV weight w (None: W = 1) in the
    lambda D^T W D term.  x starts at 0 and stays the last iterate across
    passes; the pass tv_i > 0 first sets w = tv_weight(x) (recon.cu:2022-2034);
    each pass recomputes r = b - A x; on restart 0 the pass resets the best
    residual and beta2 to |r|^2 (recon.cu:2043-2048), later restarts carry
    both; reconBest (the output, :2107-2110) is replaced by every iterate
    whose |r|^2 is below the best.  Without TV only the first pass runs.
    Returns (reconBest, per-pass residual histories)."""
    st, dot, sd = _vec_ops(vec_dtype)
    b = st(np.asarray(b).reshape(-1))
    normb = dot(b, b)
    x = st(np.zeros_like(b))
    best_x = x.copy()
    res_best = beta2 = None
    w = None
    hists = []
    for restart in range(restarts):
        for tv_i in range(tv_num if enable_tv else 1):
            if tv_i != 0:
                w = tv_weight(x.reshape(shape))
            A = make_apply(w)
            r = st(b - st(A(x)))
            if restart == 0:
                res_best = beta2 = dot(r, r)
            p = r.copy()
            hist = []
            for _ in range(num_iter):
                q = st(A(p))
                beta = beta2
                alpha = sd(beta / dot(p, q))
                x = st(x + alpha * p)
                r = st(r - alpha * q)
                beta2 = dot(r, r)
                p = st(r + sd(beta2 / beta) * p)
                hist.append(np.sqrt(beta2 / normb))
                if beta2 < res_best:
                    res_best, best_x = beta2, x.copy()
                if np.sqrt(beta2 / normb) < tol:
                    break
            hists.append(hist)
    return best_x, hists


def chan_circulant(t, dims):
    """T. Chan's optimal circulant of a multilevel Toeplitz matrix whose
    kernel t is given on the 2N grid (axis d index m = delta + N_d, delta in
    [-N_d, N_d); the standard Q-grid layout of main_mri.cu:461-500):
    per axis c_j = ((N - j) t_j + j t_{j-N}) / N, j in [0, N).  Returns the
    first column c, shaped dims.  Not in the reference (its CG is
    unpreconditioned); T. Chan, SIAM J. Sci. Stat. Comput. 9 (1988)."""
    t = np.asarray(t, np.complex128).reshape([2 * n if n > 1 else 1 for n in dims])
    c = t
    for ax, n in enumerate(dims):
        if n == 1:
            continue
        j = np.arange(n)
        shape = [1] * len(dims)
        shape[ax] = n
        w0 = ((n - j) / n).reshape(shape)
        w1 = (j / n).reshape(shape)
        pos = np.take(c, n + j, axis=ax)      # t_j
        neg = np.take(c, j, axis=ax)          # t_{j - N}
        c = w0 * pos + w1 * neg
    return c


def periodic_laplacian_eigs(dims):
    """Eigenvalues sum_d (2 - 2 cos(2 pi k_d / N_d)) of sum_d D_d^T D_d with
    the periodic first differences of fd_reg, in natural DFT order."""
    out = np.zeros(dims)
    for ax, n in enumerate(dims):
        if n == 1:
            continue
        shape = [1] * len(dims)
        shape[ax] = n
        out = out + (2 - 2 * np.cos(2 * np.pi * np.arange(n) / n)).reshape(shape)
    return out


def chan_precond_weights(q, dims, lam, floor=1e-3):
    """Weights of M^-1 = F^H diag(w) F (unnormalised DFT F, so w includes the
    1/prod(dims) of the inverse): eigenvalues of chan_circulant(sum_l Q_l)
    plus lam times the Laplacian spectrum, floored at `floor` times the DC
    one (1e-3 on the GPU; the Chan eigenvalues are Rayleigh quotients of the
    Toeplitz matrix at Fourier vectors, so a PSD kernel needs no floor)."""
    t = np.asarray(q).reshape(np.asarray(q).shape[0], -1).sum(axis=0)
    c = chan_circulant(t, dims)
    eig = np.fft.fftn(c).real + lam * periodic_laplacian_eigs(dims)
    dc = np.fft.fftn(c).real.flat[0]
    eig = np.maximum(eig, floor * abs(dc))
    return 1.0 / (eig * np.prod(dims))


def circulant_apply(w, r):
    """z = IDFT(w * DFT(r)) * prod(dims) with w from chan_precond_weights."""
    r = np.asarray(r, np.complex128).reshape(w.shape)
    return np.fft.ifftn(w * np.fft.fftn(r)) * w.size


def pcg(apply, precond, b, max_iter, tol=1e-16, best_iterate=True, on_iter=None):
    """Textbook preconditioned CG (the IMP_CG_PRECOND path of imp_toep_cg):
    rho = Re(r^H z), alpha = rho / Re(p^H A p), p = z + (rho'/rho) p; the stop
    test and best-iterate tracking use the unpreconditioned |r|.
    on_iter(it, x, rel_residual) is called after every iteration."""
    b = np.asarray(b, np.complex128).reshape(-1)
    x = np.zeros_like(b)
    r = b.copy()
    z = np.asarray(precond(r)).reshape(-1)
    p = z.copy()
    rho = np.vdot(r, z).real
    normb = np.vdot(b, b).real
    best, xbest = normb, x.copy()
    hist = []
    it = 0
    while it < max_iter:
        q = np.asarray(apply(p)).reshape(-1)
        alpha = rho / np.vdot(p, q).real
        x = x + alpha * p
        r = r - alpha * q
        rr = np.vdot(r, r).real
        z = np.asarray(precond(r)).reshape(-1)
        rho_new = np.vdot(r, z).real
        p = z + (rho_new / rho) * p
        rho = rho_new
        hist.append(np.sqrt(rr / normb))
        it += 1
        if on_iter is not None:
            on_iter(it, x, hist[-1])
        if rr < best:
            best, xbest = rr, x.copy()
        if np.sqrt(rr / normb) < tol:
            break
    return (xbest if best_iterate else x), it, hist


def bruteforce_cg(F, FH, R, y, x0, cg_num, vec_dtype=np.complex128):
    """bruteForceGpu's CG (bruteForce/bruteForceGpu.cu:470-654):
    cg_num - 1 iterations from x0 with grad = F^H(y - F x) - R x,
    d = grad + gamma d, alpha = Re(d^H grad) / (|F d|^2 + Re((R d)^H d)),
    Af += alpha F d, x += alpha d.  F, FH, R are callables (R may be None).
    vec_dtype=complex64 stores the vectors in FP32 (scalars stay FP64)."""
    st, dot, sd = _vec_ops(vec_dtype)
    y = st(y)
    x = st(x0)
    Af = st(F(x))
    d = None
    old = 0.0
    for cg_i in range(1, cg_num):
        g = st(FH(st(y - Af)))
        if R is not None:
            g = st(g - st(R(x)))
        new = dot(g, g)
        if cg_i == 1:
            d = g.copy()
        else:
            gamma = 0.0 if old == 0.0 else new / old
            d = st(g + sd(gamma) * d)
        old = new
        q = st(F(d))
        a_ctcd = dot(st(R(d)), d) if R is not None else 0.0
        alpha = sd(dot(d, g) / (dot(q, q) + a_ctcd))
        Af = st(Af + alpha * q)
        x = st(x + alpha * d)
    return x
