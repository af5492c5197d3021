"""Pins of the FP64 oracle (CPU only).

Every oracle function is checked against something that does not share its
code: closed forms, textbook library routines (numpy.fft, numpy.linalg,
an explicit sparse difference matrix), identities the reference itself tests
(the structured-vector adjoint test of bruteForceGpu.cu:240-303), and the
independent FP32 C restatement of the reference CPU loops (oracle/ref_cpu.c).
A plausible slip (sign of the phase, a dropped field-map or z term, a
transposed grid axis, a wrong Hanning centre, an off-by-one shift) fails at
least one test below.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import REPO, rel_l2


def _orc():
    from oracle import oracle
    return oracle


def _tiny(**kw):
    from impatient import synth
    return synth.tiny(**kw)


# ------------------------------------------------------------------ NUDFT
def test_single_sample_closed_form():
    orc = _orc()
    rng = np.random.default_rng(0)
    x, y, z, fm = (rng.uniform(-.5, .5, 37) for _ in range(4))
    fm = fm * 900.0
    kx, ky, kz, t = np.array([3.25]), np.array([-7.5]), np.array([2.0]), np.array([4e-3])
    d = np.array([[0.3 - 1.2j]])
    got = orc.nudft_adjoint(kx, ky, kz, t, x, y, z, fm, d)
    ph = 2 * np.pi * (3.25 * x - 7.5 * y + 2.0 * z) + fm * 4e-3
    assert np.allclose(got, d[0, 0] * np.exp(1j * ph), rtol=0, atol=1e-12)
    rho = rng.standard_normal(37) + 1j * rng.standard_normal(37)
    fwd = orc.nudft_forward(kx, ky, kz, t, x, y, z, fm, rho)
    assert np.allclose(fwd[0, 0], np.sum(rho * np.exp(-1j * ph)), atol=1e-10)


def test_cartesian_grid_is_the_dft():
    """Full Cartesian k-grid, w = 0: the forward NUDFT is numpy's 2D DFT with
    centred index conventions."""
    orc = _orc()
    N = 8
    idx = np.arange(N) - N // 2
    ky, kx = np.meshgrid(idx, idx, indexing="ij")
    iy, ix = np.meshgrid(idx / N, idx / N, indexing="ij")
    rng = np.random.default_rng(1)
    img = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    d = orc.nudft_forward(kx.ravel(), ky.ravel(), None, np.zeros(N * N), ix.ravel(), iy.ravel(),
                          None, None, img.ravel())[0].reshape(N, N)
    ref = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(img)))
    assert rel_l2(d, ref) < 1e-12
    back = orc.nudft_adjoint(kx.ravel(), ky.ravel(), None, np.zeros(N * N), ix.ravel(),
                             iy.ravel(), None, None, d.ravel()[None, :]).reshape(N, N)
    assert rel_l2(back, N * N * img) < 1e-12


def test_reference_structured_adjoint_test():
    """The adjoint test vectors of bruteForceGpu.cu:240-303 (a: (1, M) with
    every 10th entry (N, 1); b: (N+1, 1) with every 10th (1, M+2))."""
    orc = _orc()
    p = _tiny(N=16, ncoils=1)
    n, m = p.nvox, p.nk
    i = np.arange(n)
    a = np.where(i % 10 != 0, 1 + 1j * m, n + 1j)
    k = np.arange(m)
    b = np.where(k % 10 != 0, (n + 1) + 1j, 1 + 1j * (m + 2))
    o1 = np.vdot(orc.nudft_forward(p.kx, p.ky, None, p.t, p.x, p.y, None, p.fm, a)[0], b).real
    o2 = np.vdot(a, orc.nudft_adjoint(p.kx, p.ky, None, p.t, p.x, p.y, None, p.fm, b[None])).real
    assert abs(o1 - o2) / abs(o1) < 1e-12


def test_sense_adjoint_identity_3d():
    orc = _orc()
    p = _tiny(N=8, Nz=4, ncoils=3)
    rng = np.random.default_rng(2)
    a = rng.standard_normal(p.nvox) + 1j * rng.standard_normal(p.nvox)
    b = rng.standard_normal((3, p.nk)) + 1j * rng.standard_normal((3, p.nk))
    Fa = orc.nudft_forward(p.kx, p.ky, p.kz, p.t, p.x, p.y, p.z, p.fm, a, p.sens)
    FHb = orc.nudft_adjoint(p.kx, p.ky, p.kz, p.t, p.x, p.y, p.z, p.fm, b, p.sens)
    assert abs(np.vdot(b, Fa) - np.vdot(FHb, a)) / abs(np.vdot(b, Fa)) < 1e-12


def _cpu_lib():
    so = REPO / "oracle" / "liboracle_cpu.so"
    if not so.exists():
        pytest.skip("oracle/liboracle_cpu.so not built (make -C oracle)")
    return C.CDLL(str(so))


def _fp(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


def test_c_restatement_agrees_with_fp64():
    """ftCpu / iftCpu restated in FP32 C (the reference loop order and
    sinf/cosf) agree with the FP64 oracle to FP32 accuracy."""
    orc = _orc()
    lib = _cpu_lib()
    p = _tiny(N=16, ncoils=1)
    rng = np.random.default_rng(3)
    rho = (rng.standard_normal(p.nvox) + 1j * rng.standard_normal(p.nvox)).astype(np.complex64)
    keep = [_fp(v) for v in (p.kx, p.ky, p.x, p.y, p.fm, p.t, rho.real, rho.imag)]
    kr = np.zeros(p.nk, np.float32)
    ki = np.zeros(p.nk, np.float32)
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    lib.ft_cpu(P(kr), P(ki), keep[6][1], keep[7][1], keep[0][1], keep[1][1], keep[2][1],
               keep[3][1], keep[4][1], keep[5][1], C.c_int(p.nk), C.c_int(p.nvox))
    ref = orc.nudft_forward(p.kx, p.ky, None, p.t, p.x, p.y, None, p.fm, rho)[0]
    assert rel_l2(kr + 1j * kr * 0 + 1j * ki, ref) < 1e-4
    ir = np.zeros(p.nvox, np.float32)
    ii = np.zeros(p.nvox, np.float32)
    d = ref.astype(np.complex64)
    dd = [_fp(d.real), _fp(d.imag)]
    lib.ift_cpu(P(ir), P(ii), dd[0][1], dd[1][1], keep[0][1], keep[1][1], keep[2][1],
                keep[3][1], keep[4][1], keep[5][1], C.c_int(p.nk), C.c_int(p.nvox))
    ref2 = orc.nudft_adjoint(p.kx, p.ky, None, p.t, p.x, p.y, None, p.fm, d[None])
    assert rel_l2(ir + 1j * ii, ref2) < 1e-4


def test_c_bruteforce_restatement_agrees_with_fp64():
    """bruteForceCpu's CG restated in FP32 C (bruteforce_cpu: the timed CPU
    brute-force baseline) follows the FP64 oracle CG (with the same finite
    differences) on a 2-coil 12x12 problem: 5 iterations from zero."""
    orc = _orc()
    lib = _cpu_lib()
    p = _tiny(N=12, ncoils=2)
    N, nc, lam, cg = 12, 2, 5.0, 6
    rng = np.random.default_rng(9)
    y = (rng.standard_normal((nc, p.nk)) + 1j * rng.standard_normal((nc, p.nk))).astype(np.complex64)
    F = lambda v: orc.nudft_forward(p.kx, p.ky, None, p.t, p.x, p.y, None, p.fm, v, p.sens)
    FH = lambda d: orc.nudft_adjoint(p.kx, p.ky, None, p.t, p.x, p.y, None, p.fm,
                                     np.asarray(d).reshape(nc, -1), p.sens)
    R = lambda v: orc.fd_reg(np.asarray(v).reshape(N, N), lam).reshape(-1)
    x0 = np.zeros(p.nvox, np.complex64)
    ref = orc.bruteforce_cg(F, FH, R, y, x0, cg)
    keep = [_fp(v) for v in (p.kx, p.ky, p.x, p.y, p.fm, p.t, y.real.reshape(-1),
                            y.imag.reshape(-1), p.sens.real.reshape(-1), p.sens.imag.reshape(-1))]
    xr = np.zeros(p.nvox, np.float32)
    xi = np.zeros(p.nvox, np.float32)
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    lib.bruteforce_cpu(P(xr), P(xi), keep[6][1], keep[7][1], keep[0][1], keep[1][1], keep[2][1],
                       keep[3][1], keep[4][1], keep[5][1], keep[8][1], keep[9][1], C.c_int(nc),
                       C.c_int(p.nk), C.c_int(N), C.c_int(N), C.c_int(cg), C.c_float(lam))
    err = rel_l2(xr + 1j * xi, ref)
    assert err < 1e-3, err


# ---------------------------------------------------------------------- Q
def test_hanning_partition_of_unity():
    orc = _orc()
    t = np.linspace(1e-3, 9e-3, 777)
    for L in (1, 3, 8):
        tau, t0 = (t[-1] - t[0]) / L, t[0]
        s = sum(orc.hanning(t, tau, l, t0) for l in range(L + 1))
        assert np.allclose(s, 1.0, atol=1e-12)
        # centre of segment l sits at t0 + l tau
        assert orc.hanning(t0 + 2 * tau if L >= 2 else t0, tau, 2 if L >= 2 else 0, t0) == 1.0


def test_q_properties():
    orc = _orc()
    p = _tiny(N=8, ncoils=1, ntime_segs=3)
    q = orc.toeplitz_q(p.kx, p.ky, None, p.t, 8, 8, 1, 3)
    G = q.reshape(4, 16, 16)
    # Hermitian: Q(-r) = conj(Q(r)) on the grid r in [-N, N) (index i <-> r = i - N)
    inner = G[:, 1:, 1:]
    flipped = G[:, 1:, 1:][:, ::-1, ::-1]
    assert np.allclose(inner, np.conj(flipped), atol=1e-9)
    # sum over segments = the unsegmented kernel sum_k e^{i 2pi k.r}
    gx, gy, _ = orc.q_grid(8, 8, 1)
    full = np.exp(1j * 2 * np.pi * (np.outer(p.kx.astype(float), gx) +
                                   np.outer(p.ky.astype(float), gy))).sum(0)
    tau, t0 = orc.segment_timing(p.t, 3)
    inside = np.ones(p.nk, bool)
    assert rel_l2(q.sum(0), full) < 1e-6
    # r = 0 holds the total segment weight
    w0 = [orc.hanning(p.t, tau, l, t0).sum() for l in range(4)]
    assert np.allclose(G[:, 8, 8].real, w0) and np.allclose(G[:, 8, 8].imag, 0, atol=1e-9)
    assert inside.all()
    # cumulative (reference direct path) = prefix sums
    qc = orc.toeplitz_q(p.kx, p.ky, None, p.t, 8, 8, 1, 3, cumulative=True)
    assert np.allclose(qc, np.cumsum(q, axis=0))


def test_q_grid_layout():
    """main_mri.cu:461-500: y outer, x inner (x fastest), r = x / Nx."""
    orc = _orc()
    gx, gy, gz = orc.q_grid(4, 3, 1)
    assert gx[:8].tolist() == [(-4 + i) / 4 for i in range(8)]
    assert np.all(gy[:8] == -1.0) and gy[8] == -2 / 3
    gx3, gy3, gz3 = orc.q_grid(2, 2, 2)
    assert gz3[:4].tolist() == [-1.0, -0.5, 0.0, 0.5] and np.all(gx3[:4] == -1.0)


def test_c_q_restatement_agrees_with_fp64():
    orc = _orc()
    lib = _cpu_lib()
    p = _tiny(N=8, ncoils=1, ntime_segs=2)
    gx, gy, _ = orc.q_grid(8, 8, 1)
    tau, t0 = orc.segment_timing(p.t, 2)
    ref = orc.toeplitz_q(p.kx, p.ky, None, p.t, 8, 8, 1, 2)
    keep = [_fp(v) for v in (p.kx, p.ky, p.t, gx, gy)]
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    for l in range(3):
        qr = np.zeros(len(gx), np.float32)
        qi = np.zeros(len(gx), np.float32)
        lib.q_cpu(P(qr), P(qi), keep[0][1], keep[1][1], None, keep[2][1], None, keep[3][1],
                  keep[4][1], None, C.c_int(p.nk), C.c_int(len(gx)), C.c_float(l),
                  C.c_float(tau), C.c_float(t0))
        assert rel_l2(qr + 1j * qi, ref[l]) < 1e-4


# --------------------------------------------------------------- Toeplitz
@pytest.mark.parametrize("Nz", [1, 4])
def test_toeplitz_without_fieldmap_is_normal_operator(Nz):
    """w = 0: sum_l T_l = F^H F exactly (circulant embedding, partition of unity)."""
    orc = _orc()
    p = _tiny(N=8, Nz=Nz, ncoils=2, ntime_segs=2, max_hz=0.0)
    fm0 = np.zeros(p.nvox)
    q = orc.toeplitz_q(p.kx, p.ky, p.kz, p.t, 8, 8, Nz, 2)
    tau, t0 = orc.segment_timing(p.t, 2)
    rng = np.random.default_rng(4)
    v = rng.standard_normal(p.nvox) + 1j * rng.standard_normal(p.nvox)
    shape = (8, 8, Nz) if Nz > 1 else (8, 8)
    T = orc.toeplitz_apply(v.reshape(shape), q, tau, t0, fm0, p.sens).reshape(-1)
    F = orc.nudft_forward(p.kx, p.ky, p.kz, p.t, p.x, p.y, p.z, fm0, v, p.sens)
    FHF = orc.nudft_adjoint(p.kx, p.ky, p.kz, p.t, p.x, p.y, p.z, fm0, F, p.sens)
    assert rel_l2(T, FHF) < 1e-12


def test_toeplitz_matches_segmented_direct_formula():
    """w != 0: the apply equals sum_c S_c^H sum_l e^{+iw tau_l} F^H diag(a_l) F
    e^{-iw tau_l} S_c, evaluated by direct sums."""
    orc = _orc()
    L = 3
    p = _tiny(N=8, ncoils=2, ntime_segs=L, max_hz=80.0)
    q = orc.toeplitz_q(p.kx, p.ky, None, p.t, 8, 8, 1, L)
    tau, t0 = orc.segment_timing(p.t, L)
    rng = np.random.default_rng(5)
    v = rng.standard_normal(p.nvox) + 1j * rng.standard_normal(p.nvox)
    T = orc.toeplitz_apply(v.reshape(8, 8), q, tau, t0, p.fm, p.sens).reshape(-1)
    fm = p.fm.astype(np.float64)
    acc = np.zeros(p.nvox, complex)
    for c in range(2):
        for l in range(L + 1):
            a = orc.hanning(p.t, tau, l, t0)
            ph = np.exp(-1j * fm * (tau * l + t0))
            u = ph * p.sens[c] * v
            Fu = orc.nudft_forward(p.kx, p.ky, None, p.t, p.x, p.y, None, None, u)
            acc += np.conj(p.sens[c]) * np.conj(ph) * orc.nudft_adjoint(
                p.kx, p.ky, None, p.t, p.x, p.y, None, None, a * Fu)
    assert rel_l2(T, acc) < 1e-10


# ----------------------------------------------------------- regularizer
def _periodic_D(n):
    """First difference with periodic boundary, (Dp)(i) = p(i) - p(i+1 mod n),
    as createDWithPeriodicBoundary.m builds it."""
    D = np.eye(n)
    D[np.arange(n), (np.arange(n) + 1) % n] -= 1
    return D


@pytest.mark.parametrize("shape", [(6, 5), (4, 3, 5)])
def test_fd_reg_equals_explicit_matrix(shape):
    orc = _orc()
    n = int(np.prod(shape))
    rng = np.random.default_rng(6)
    w = rng.uniform(0.2, 2.0, n)
    blocks = []
    for ax, na in enumerate(shape):
        mats = [np.eye(s) for s in shape]
        mats[ax] = _periodic_D(na)
        K = mats[0]
        for m_ in mats[1:]:
            K = np.kron(K, m_)
        blocks.append(K)
    R = sum(B.T @ np.diag(w) @ B for B in blocks)
    p = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel_l2(orc.fd_reg(p.reshape(shape), 0.7, w.reshape(shape)).reshape(-1),
                  0.7 * R @ p) < 1e-13


def _difference_coo(shape):
    """C = [D_0; D_1; ...] (one block of periodic first differences per image
    axis, the matrix createDWithPeriodicBoundary.m writes as c.rmtx) as
    0-based COO."""
    n = int(np.prod(shape))
    I, J, V = [], [], []
    for b, na in enumerate(shape):
        mats = [np.eye(s) for s in shape]
        mats[b] = _periodic_D(na)
        K = mats[0]
        for m_ in mats[1:]:
            K = np.kron(K, m_)
        r, c = np.nonzero(K)
        I.append(r + b * n)
        J.append(c)
        V.append(K[r, c])
    return len(shape) * n, n, np.concatenate(I), np.concatenate(J), np.concatenate(V)


@pytest.mark.parametrize("shape", [(6, 5), (4, 3, 5)])
def test_sparse_reg_with_difference_matrix_equals_fd_reg(shape):
    """-reg with C = the periodic difference matrix is the -fd stencil."""
    orc = _orc()
    rng = np.random.default_rng(8)
    rows, cols, I, J, V = _difference_coo(shape)
    w = rng.uniform(0.2, 2.0, shape)
    p = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    got = orc.sparse_reg(p, 0.7, rows, cols, I, J, V, w)
    assert rel_l2(got, orc.fd_reg(p, 0.7, w)) < 1e-13


def test_sparse_reg_dense_brute_force():
    """Random C with duplicate entries: lam C^T diag([w; w]) C p densely."""
    orc = _orc()
    rng = np.random.default_rng(9)
    n, rows, nnz = 20, 40, 90
    I = rng.integers(0, rows, nnz)
    J = rng.integers(0, n, nnz)
    I[:5], J[:5] = I[5:10], J[5:10]          # duplicates add
    V = rng.standard_normal(nnz)
    Cd = np.zeros((rows, n))
    np.add.at(Cd, (I, J), V)
    w = rng.uniform(0.5, 1.5, n)
    p = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = 0.3 * Cd.T @ (np.concatenate([w, w]) * (Cd @ p))
    assert rel_l2(orc.sparse_reg(p, 0.3, rows, n, I, J, V, w), ref) < 1e-13
    assert rel_l2(orc.sparse_reg(p, 0.3, rows, n, I, J, V),
                  0.3 * Cd.T @ (Cd @ p)) < 1e-13


def test_fd_reg_plane_wave_eigenvalue():
    orc = _orc()
    N1, N2 = 8, 12
    f1, f2 = 3, 5
    y, x = np.meshgrid(np.arange(N1), np.arange(N2), indexing="ij")
    v = np.exp(2j * np.pi * (f1 * y / N1 + f2 * x / N2))
    lam = 0.3
    eig = lam * ((2 - 2 * np.cos(2 * np.pi * f1 / N1)) + (2 - 2 * np.cos(2 * np.pi * f2 / N2)))
    assert rel_l2(orc.fd_reg(v, lam), eig * v) < 1e-13


def test_tv_weight_closed_form():
    orc = _orc()
    x = np.zeros((4, 4), complex)
    x[1, 2] = 3.0
    w = orc.tv_weight(x)
    # at (1,2): |x - right|^2 + |x - below|^2 = 9 + 9
    assert np.isclose(w[1, 2], 0.5 / np.sqrt(18 + 1e-6))
    assert np.isclose(w[0, 0], 0.5 / np.sqrt(1e-6))


# --------------------------------------------------------------------- CG
def _spd(n, seed):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return B @ B.conj().T + n * np.eye(n)


def test_toeplitz_cg_solves_spd_system():
    orc = _orc()
    A = _spd(20, 7)
    b = np.arange(20) + 1j
    x, it, hist = orc.toeplitz_cg(lambda v: A @ v, b, 200, tol=1e-13)
    assert rel_l2(x, np.linalg.solve(A, b)) < 1e-10
    assert it <= 40


def test_toeplitz_cg_best_iterate():
    """reconBest: the returned iterate has the smallest recursive residual."""
    orc = _orc()
    A = _spd(30, 8)
    b = np.ones(30, complex)
    xb, it, hist = orc.toeplitz_cg(lambda v: A @ v, b, 6, best_iterate=True)
    xl, _, _ = orc.toeplitz_cg(lambda v: A @ v, b, 6, best_iterate=False)
    k = int(np.argmin(hist)) + 1
    xk, _, _ = orc.toeplitz_cg(lambda v: A @ v, b, k, best_iterate=False)
    assert np.allclose(xb, xk)
    if k == 6:
        assert np.allclose(xb, xl)


def test_bruteforce_cg_converges_to_normal_equations():
    orc = _orc()
    rng = np.random.default_rng(9)
    Fm = rng.standard_normal((40, 12)) + 1j * rng.standard_normal((40, 12))
    Rm = _spd(12, 10) * 0.01
    y = rng.standard_normal(40) + 1j * rng.standard_normal(40)
    x = orc.bruteforce_cg(lambda v: Fm @ v, lambda d: Fm.conj().T @ d, lambda v: Rm @ v, y,
                          np.zeros(12, complex), 40)
    xs = np.linalg.solve(Fm.conj().T @ Fm + Rm, Fm.conj().T @ y)
    assert rel_l2(x, xs) < 1e-8
    # FP32 vector storage stays close on this well-conditioned system
    x32 = orc.bruteforce_cg(lambda v: Fm @ v, lambda d: Fm.conj().T @ d, lambda v: Rm @ v, y,
                            np.zeros(12, complex), 40, vec_dtype=np.complex64)
    assert rel_l2(x32, xs) < 1e-4


# ------------------------------------------------------- preconditioner
def _toeplitz_matrix(t, dims):
    """Dense multilevel Toeplitz matrix T[i, j] = t(r_i - r_j) from a kernel
    on the 2N grid (index delta + N per axis)."""
    grids = np.meshgrid(*[np.arange(n) for n in dims], indexing="ij")
    pts = np.stack([g.reshape(-1) for g in grids], axis=1)
    tt = np.asarray(t).reshape([2 * n if n > 1 else 1 for n in dims])
    d = pts[:, None, :] - pts[None, :, :]
    idx = tuple((d[..., a] + (n if n > 1 else 0)) for a, n in enumerate(dims))
    return tt[idx]


def test_chan_circulant_is_frobenius_optimal():
    """chan_circulant equals the circulant nearest to T in Frobenius norm,
    found here by brute force: the least-squares fit over all circulants
    (mean of each wrapped multilevel diagonal)."""
    orc = _orc()
    rng = np.random.default_rng(21)
    for dims in [(6,), (4, 5), (3, 4, 2)]:
        t = rng.standard_normal([2 * n for n in dims]) + 1j * rng.standard_normal(
            [2 * n for n in dims])
        T = _toeplitz_matrix(t, dims)
        n = int(np.prod(dims))
        grids = np.meshgrid(*[np.arange(k) for k in dims], indexing="ij")
        pts = np.stack([g.reshape(-1) for g in grids], axis=1)
        # circulant class of entry (i, j): (r_i - r_j) mod dims
        cls = np.ravel_multi_index(tuple(((pts[:, None, a] - pts[None, :, a]) % dims[a])
                                         for a in range(len(dims))), dims)
        best = np.array([T[cls == k].mean() for k in range(n)]).reshape(dims)
        got = orc.chan_circulant(t, dims)
        assert rel_l2(got, best) < 1e-12


def test_chan_weights_invert_a_circulant_exactly():
    """When the Toeplitz kernel is N-periodic-compatible (t(delta) = t(delta
    - N) for every wrap) the Chan circulant equals T's circulant extension,
    and M^-1 (T + lam R) v = v for the circulant part: check against a dense
    solve of C + lam L with L the explicit periodic difference Laplacian."""
    orc = _orc()
    dims = (8, 6)
    rng = np.random.default_rng(22)
    c = rng.standard_normal(dims) + 1j * rng.standard_normal(dims)
    # Hermitian positive circulant: eigenvalues > 0
    eig = rng.uniform(1, 3, dims)
    c = np.fft.ifftn(eig)
    t = np.zeros([2 * n for n in dims], complex)
    for a in range(2 * dims[0]):
        for b in range(2 * dims[1]):
            t[a, b] = c[(a - dims[0]) % dims[0], (b - dims[1]) % dims[1]]
    lam = 0.2
    w = orc.chan_precond_weights(t[None], dims, lam)
    v = rng.standard_normal(dims) + 1j * rng.standard_normal(dims)
    C_v = np.fft.ifftn(eig * np.fft.fftn(v))
    Av = C_v + orc.fd_reg(v, lam)
    assert rel_l2(orc.circulant_apply(w, Av), v) < 1e-12


def test_pcg_matches_solve_and_needs_fewer_iterations():
    """PCG with the Chan circulant on a variable-density Toeplitz system
    (radially weighted samples) solves it and converges in fewer iterations
    than plain CG."""
    orc = _orc()
    N = 12
    rng = np.random.default_rng(23)
    M = 3000
    r = np.sqrt(rng.uniform(0, 1, M)) ** 2 * (N / 2)   # dense at the centre
    th = rng.uniform(0, 2 * np.pi, M)
    kx, ky = r * np.cos(th), r * np.sin(th)
    t = np.linspace(0, 1e-3, M)
    q = orc.toeplitz_q(kx, ky, None, t, N, N, 1, 1)   # two segments, a_0 + a_1 = 1
    tau, t0 = orc.segment_timing(t, 1)
    lam = 1e-2 * M / N ** 2
    A = lambda v: orc.toeplitz_apply(v.reshape(N, N), q, tau, t0, None, None,
                                     lam).reshape(-1)
    w = orc.chan_precond_weights(q, (N, N), lam)
    Mi = lambda v: orc.circulant_apply(w, v).reshape(-1)
    b = rng.standard_normal(N * N) + 1j * rng.standard_normal(N * N)
    n = N * N
    Ad = np.stack([A(e) for e in np.eye(n)], axis=1)
    xs = np.linalg.solve(Ad, b)
    xp, itp, hp = orc.pcg(A, Mi, b, 500, tol=1e-10)
    xc, itc, hc = orc.toeplitz_cg(A, b, 500, tol=1e-10)
    assert rel_l2(xp, xs) < 1e-8 and rel_l2(xc, xs) < 1e-8
    assert itp < 0.7 * itc, (itp, itc)


# --------------------------------------------------------- learned start
def test_cnn_oracle_backward_matches_finite_differences():
    """The oracle's backward pass (FP64, no bf16 rounding) is the derivative
    of its forward loss: central differences on a sample of parameters from
    every layer, weights and biases."""
    from oracle import cnn
    rng = np.random.default_rng(31)
    p = cnn.random_params(3)
    z = rng.standard_normal((5, 7)) + 1j * rng.standard_normal((5, 7))
    t = z + 0.3 * (rng.standard_normal((5, 7)) + 1j * rng.standard_normal((5, 7)))
    loss, g = cnn.loss_and_grad(p, z, t, bf16_round=False)
    idx = np.flatnonzero(cnn.trainable_mask())
    pick = rng.choice(idx, 40, replace=False)
    pick = np.concatenate([pick, [l * cnn.PSZ + cnn.WSZ for l in range(cnn.NLAYER)]])
    h = 1e-6
    for i in pick:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = cnn.loss_and_grad(pp, z, t, bf16_round=False)
        lm, _ = cnn.loss_and_grad(pm, z, t, bf16_round=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-7, (i, fd, g[i])


def test_cnn_oracle_conv_is_correlation_and_adjoint():
    """conv3x3 against a direct per-pixel loop, and <conv(a), g> = <a, bwd(g)>
    (the backward-data pass is the adjoint of the convolution)."""
    from oracle import cnn
    rng = np.random.default_rng(32)
    a = rng.standard_normal((4, 6, cnn.CH))
    W = rng.standard_normal((9, cnn.CH, cnn.CH))
    out = cnn.conv3x3(a, W)
    y, x, co = 0, 5, 3
    ref = 0.0
    for tap in range(9):
        yy, xx = y + tap // 3 - 1, x + tap % 3 - 1
        if 0 <= yy < 4 and 0 <= xx < 6:
            ref += a[yy, xx] @ W[tap, co]
    assert np.isclose(out[y, x, co], ref)
    g = rng.standard_normal(out.shape)
    assert np.isclose(np.sum(out * g), np.sum(a * cnn.conv3x3_bwd_data(g, W)))
    dW, _ = cnn.wgrad(a, g)
    # <conv(a; W), g> is linear in W with gradient dW
    assert np.isclose(np.sum(out * g), np.sum(W * dW))


def test_bf16_rounding_is_nearest_even():
    from oracle import cnn
    one = np.float32(1.0)
    ulp = 2.0 ** -7
    assert cnn.bf16(np.float32(1.0 + ulp / 2)) == 1.0          # tie -> even
    assert cnn.bf16(np.float32(1.0 + 3 * ulp / 2)) == 1.0 + 2 * ulp
    assert cnn.bf16(np.float32(1.0 + ulp / 2 + 2 ** -20)) == 1.0 + ulp
    assert cnn.bf16(one) == 1.0 and cnn.bf16(np.float32(-3.0)) == -3.0


@pytest.mark.parametrize("shape", [(8, 8, 1), (4, 6, 3)])
def test_toeplitz_operator_equals_apply(shape):
    """The cached operator (Qf once, scipy.fft) is the same map as
    toeplitz_apply, in 2D and 3D, with coils, field map and regularizer."""
    orc = _orc()
    rng = np.random.default_rng(21)
    N1, N2, N3 = shape
    d3 = N3 > 1
    M = (2 * N1, 2 * N2) + ((2 * N3,) if d3 else ())
    L = 2
    q = rng.standard_normal((L + 1, int(np.prod(M)))) + 1j * rng.standard_normal((L + 1, int(np.prod(M))))
    img = shape if d3 else shape[:2]
    fm = rng.uniform(-300, 300, img)
    sens = rng.standard_normal((2,) + img) + 1j * rng.standard_normal((2,) + img)
    v = rng.standard_normal(img) + 1j * rng.standard_normal(img)
    ref = orc.toeplitz_apply(v, q, 1e-3, 2e-4, fm, sens, 0.3)
    A = orc.toeplitz_operator(q, 1e-3, 2e-4, shape, fm, sens, 0.3)
    got = A(v.reshape(-1)).reshape(img)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-12
    # a second call on another vector: the operator keeps no state between calls
    v2 = rng.standard_normal(img) + 1j * rng.standard_normal(img)
    ref2 = orc.toeplitz_apply(v2, q, 1e-3, 2e-4, fm, sens, 0.3)
    got2 = A(v2.reshape(-1)).reshape(img)
    assert np.linalg.norm(got2 - ref2) / np.linalg.norm(ref2) < 1e-12


def _dense_fd(shape, w):
    """Dense sum_d D_d^T diag(w) D_d with periodic first differences."""
    blocks = []
    for ax, na in enumerate(shape):
        if na == 1:
            continue
        mats = [np.eye(s) for s in shape]
        mats[ax] = _periodic_D(na)
        K = mats[0]
        for m_ in mats[1:]:
            K = np.kron(K, m_)
        blocks.append(K)
    return sum(B.T @ np.diag(np.asarray(w).reshape(-1)) @ B for B in blocks)


@pytest.mark.parametrize("shape", [(5, 6), (3, 4, 5)])
def test_toeplitz_recon_tv_passes_solve_reweighted_systems(shape):
    """Converged TV passes of toeplitz_recon: pass k solves
    (G + lam D^T W(x_{k-1}) D) x_k = b with W = 0.5/sqrt(sum |D x|^2 + 1e-6)
    of the previous pass's image (W = 1 in pass 0), checked against dense
    solves; dropping the reweighting changes the answer."""
    orc = _orc()
    n = int(np.prod(shape))
    G = _spd(n, 31)
    lam = 40.0
    rng = np.random.default_rng(32)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)

    def make_apply(w):
        return lambda v: G @ v + orc.fd_reg(np.asarray(v).reshape(shape), lam, w).reshape(-1)
    x, hists = orc.toeplitz_recon(make_apply, b, shape, 400, enable_tv=True, tv_num=3,
                                  tol=1e-14)
    assert len(hists) == 3
    xk = np.linalg.solve(G + lam * _dense_fd(shape, np.ones(n)), b)
    for _ in range(2):
        w = orc.tv_weight(xk.reshape(shape))
        xk = np.linalg.solve(G + lam * _dense_fd(shape, w), b)
    x1 = np.linalg.solve(G + lam * _dense_fd(shape, np.ones(n)), b)
    assert rel_l2(x, xk) < 1e-9
    assert rel_l2(x1, xk) > 1e-3
    # without TV only one pass runs and it is the plain CG of toeplitz_cg
    x0, h0 = orc.toeplitz_recon(make_apply, b, shape, 7, enable_tv=False, tv_num=3)
    xc, _, _ = orc.toeplitz_cg(make_apply(None), b, 7)
    assert len(h0) == 1 and rel_l2(x0, xc) < 1e-12


def test_toeplitz_recon_restarts_carry_the_best_residual():
    """Restarts after the first keep reconBest unless an iterate beats every
    earlier one: the returned image has the smallest residual of the whole run
    (restart 0's passes each start their own best, recon.cu:2043-2048)."""
    orc = _orc()
    n = 24
    G = _spd(n, 33)
    b = np.linspace(1, 2, n) + 0.5j
    A = lambda v: G @ v
    x, hists = orc.toeplitz_recon(lambda w: A, b, (4, 6), 3, restarts=3)
    # candidates: the zero start (relative residual 1) and every iterate
    allres = np.concatenate([[1.0]] + [np.asarray(h) for h in hists])
    true = np.linalg.norm(b - A(x)) / np.linalg.norm(b)
    assert len(hists) == 3
    assert abs(true - allres.min()) < 1e-9 * allres.min()


def _interior_samples(rng, n, M, dims):
    """Samples at least 3 grid cells inside the k-space window on every axis."""
    return np.stack([rng.uniform(-n[a] / 2 + 3, n[a] / 2 - 3, M) for a in range(dims)], axis=1)


@pytest.mark.parametrize("img,oss", [((20, 24, 1), (1.25, 1.5, 2.0)), ((16, 12, 8), (1.5, 2.0))])
def test_gridding_is_the_adjoint_nudft_away_from_the_grid_edge(img, oss):
    """Gridding (KB width 4, Beatty beta) of samples away from the edge of
    the oversampled grid is the adjoint NUDFT on x = (i - n/2)/n: to 1e-6 at
    os = 2 and to a few 1e-3 at os = 1.5, in 2D and 3D; the error falls with
    os.  A sign, index, shift or scale mistake gives O(1) errors."""
    orc = _orc()
    rng = np.random.default_rng(51)
    nx, ny, nz = img
    dims = 3 if nz > 1 else 2
    k = _interior_samples(rng, (nx, ny, nz), 700, dims)
    d = rng.standard_normal(700) + 1j * rng.standard_normal(700)
    iy, ix, iz = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
    xs = [((ix - nx // 2) / nx).ravel(), ((iy - ny // 2) / ny).ravel(), ((iz - nz // 2) / nz).ravel()]
    ph = 2 * np.pi * sum(np.outer(k[:, a], xs[a]) for a in range(dims))
    exact = (d @ np.exp(1j * ph)).reshape(ny, nx, nz) if dims == 3 else (d @ np.exp(1j * ph)).reshape(ny, nx)
    errs = [rel_l2(orc.gridding_adjoint(k, d, img, os_), exact) for os_ in oss]
    assert errs[-1] < 1e-6
    assert errs[-2] < 1e-2
    assert all(a > b for a, b in zip(errs, errs[1:]))


def test_gridding_grid_edge_is_not_wrapped():
    """As the reference (no periodic wrap: window indices clamped to the
    grid, CPU_kernels.cpp:137-141, GPU_kernels.cu:211-221), a sample at
    k = -n/2 keeps only the part of its kernel inside the grid: the image
    centre is well below the exact 1 (0.72 at os = 2), which a wrapped grid
    would reproduce; one cell further in, the whole kernel fits."""
    orc = _orc()
    n = 32
    k = np.array([[-n / 2, 0.0]])
    g = orc.gridding_adjoint(k, np.ones(1), (n, n, 1), 2.0)
    ratio = np.abs(g[n // 2, n // 2])       # exact value 1 at x = 0
    assert 0.55 < ratio < 0.85
    g2 = orc.gridding_adjoint(np.array([[-n / 2 + 2, 0.0]]), np.ones(1), (n, n, 1), 2.0)
    assert abs(np.abs(g2[n // 2, n // 2]) - 1.0) < 1e-6


def test_gridding_params_follow_the_reference_rounding():
    """gridOS rounding and grid sizes of computeQ_GPU_Grid / computeFH_GPU_Grid:
    2D grids even (os rounded on x), 3D z a multiple of 4 (os rounded on z),
    Beatty's beta."""
    orc = _orc()
    os_, beta, G = orc.gridding_params((1024, 1024, 1), 1.375)
    assert (os_, G) == (1.375, (1408, 1408, 1))
    assert abs(beta - np.pi * np.sqrt(0.875 ** 2 * 64 / 1.375 ** 2 - 0.8)) < 1e-12
    os3, _, G3 = orc.gridding_params((128, 128, 32), 1.375)
    assert os3 == 1.375 and G3 == (176, 176, 44)
    os5, _, G5 = orc.gridding_params((128, 128, 12), 1.375)     # z rounded up to 20
    assert abs(os5 - 20 / 12) < 1e-6 and G5[2] == 20 and G5[0] % 2 == 0
    os4, _, G4 = orc.gridding_params((64, 64, 1), 1.3)
    assert G4[0] % 2 == 0 and abs(os4 * 64 - G4[0]) < 1e-4


def test_gridding_q_and_fhd_against_direct():
    """Q_l and the time-segmented F^H d by gridding against the direct sums
    on a spiral (samples out to |k| = N/2, so the unwrapped edge costs a
    little): the reference's 'around 1%' at its default os."""
    orc = _orc()
    import sys
    sys.path.insert(0, str(REPO / "impatient-mri_amd"))
    from impatient import synth
    p = synth.tiny(N=32, ncoils=2, J=8, S=256, ntime_segs=3)
    q = orc.toeplitz_q(p.kx, p.ky, None, p.t, 32, 32, 1, 3)
    qg = orc.gridding_q(p.kx, p.ky, None, p.t, 32, 32, 1, 3, 1.375)
    assert rel_l2(qg, q) < 0.02
    rng = np.random.default_rng(52)
    d = rng.standard_normal((2, p.nk)) + 1j * rng.standard_normal((2, p.nk))
    tau, t0 = orc.segment_timing(p.t, 3)
    # with the field map the exact reference is the time-segmented sum
    ts = sum(np.exp(1j * p.fm.astype(np.float64) * (tau * l + t0)) *
             orc.nudft_adjoint(p.kx, p.ky, None, p.t, p.x, p.y, None, None,
                               d * orc.hanning(p.t, tau, l, t0)[None, :], p.sens)
             for l in range(4))
    bg = orc.gridding_fhd(p.kx, p.ky, None, p.t, p.fm, d, p.sens, 32, 32, 1, 3, 2.0)
    assert rel_l2(bg, ts) < 0.05


@pytest.mark.parametrize("img,kscale", [((32, 32, 1), 1.0), ((16, 16, 8), 2.0)])
def test_c_gold_gridding_agrees_with_fp64(img, kscale):
    """gridding_Gold_2D / 3D restated in FP32 C (oracle/ref_cpu.c
    gridding_gold, the timed CPU gridding) against the FP64 grid of
    oracle.gridding_grid with the gold path's beta = 18.5547, in the
    reference's output order ([x][y] in 2D, [y][x][z] in 3D), with Hanning
    weights of one segment."""
    orc = _orc()
    lib = _cpu_lib()
    rng = np.random.default_rng(53)
    nx, ny, nz = img
    d3 = nz > 1
    M = 500
    lim = [n / (2 * kscale) for n in img]
    k = np.stack([rng.uniform(-lim[a], lim[a], M) for a in range(3 if d3 else 2)], axis=1)
    t = np.linspace(0.0, 5e-3, M)
    v = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    L, l = 3, 1
    tau, t0 = orc.segment_timing(t, L)
    os_, _, G = orc.gridding_params(img, 1.5)
    ref = orc.gridding_grid(kscale * k, v * orc.hanning(t.astype(np.float32), tau, l, t0), img,
                            1.5, beta=18.5547)
    ref = np.transpose(ref, (1, 0, 2)).reshape(-1) if d3 else ref.reshape(-1)
    keep = [_fp(a) for a in (k[:, 0], k[:, 1], k[:, 2] if d3 else np.zeros(M), v.real, v.imag, t)]
    gr = np.zeros(int(np.prod(G)), np.float32)
    gi = np.zeros(int(np.prod(G)), np.float32)
    g = (C.c_int * 3)(*G)
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    lib.gridding_gold(C.c_int(M), keep[0][1], keep[1][1], keep[2][1], C.c_float(kscale),
                      keep[3][1], keep[4][1], keep[5][1], C.c_float(l), C.c_float(tau), C.c_int(nx),
                      C.c_int(ny), C.c_int(nz), C.c_float(os_), C.c_float(18.5547), g, P(gr), P(gi))
    assert rel_l2(gr + 1j * gi, ref) < 2e-5
