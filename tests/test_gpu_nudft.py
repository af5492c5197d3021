"""GPU parity of the direct NUDFT operators against the FP64 oracle.

Tolerances (relative L2 against FP64): the operators are held to 2e-6
(split-precision phase, compensated sample sums), well inside the 1e-4
bound set for reconstructed images.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

OP_TOL = 2e-6


def _prob(**kw):
    from impatient import synth
    return synth.tiny(**kw)


@pytest.mark.parametrize("N,Nz,ncoils", [(16, 1, 1), (16, 1, 3), (12, 1, 5), (8, 4, 2)])
def test_adjoint_matches_oracle(ctx, N, Nz, ncoils):
    from impatient import ops
    from oracle import oracle as orc
    p = _prob(N=N, Nz=Nz, ncoils=ncoils)
    rng = np.random.default_rng(1)
    d = (rng.standard_normal((ncoils, p.nk)) + 1j * rng.standard_normal((ncoils, p.nk))).astype(np.complex64)
    phi = (1 + 0.3 * rng.standard_normal((ncoils, p.nk)) + 0.2j * rng.standard_normal((ncoils, p.nk)))
    got = ops.nudft_adjoint(ctx, p.kx, p.ky, p.kz, p.t, p.x, p.y, p.z, p.fm, d, p.sens, phi)
    ref = orc.nudft_adjoint(p.kx, p.ky, p.kz, p.t, p.x, p.y, p.z, p.fm, d, p.sens, phi)
    err = rel_l2(got, ref)
    print(f"adjoint N={N} Nz={Nz} C={ncoils}: rel_l2={err:.2e}")
    assert err < OP_TOL


@pytest.mark.parametrize("N,Nz,ncoils,perturb", [(8, 8, 3, False), (8, 8, 3, True),
                                                 (8, 16, 3, False), (12, 16, 4, False)])
def test_adjoint_stacked_trajectory_path(ctx, N, Nz, ncoils, perturb):
    """A stack of 2D trajectories (every kz partition repeats partition 0's
    kx, ky and t) takes the z-separable path (k_nudft_stack: 24, 48 and 64
    right-hand sides, one and two groups); one perturbed sample sends the
    same data to the direct 3D kernel.  Both match the FP64 oracle."""
    import ctypes as C
    from impatient import ops
    from impatient._lib import call
    from oracle import oracle as orc
    p = _prob(N=N, Nz=Nz, ncoils=ncoils)
    kx = p.kx.copy()
    if perturb:
        kx[len(kx) // 2 + 3] += np.float32(1e-3)
    rng = np.random.default_rng(11)
    d = (rng.standard_normal((ncoils, p.nk)) + 1j * rng.standard_normal((ncoils, p.nk))).astype(np.complex64)
    phi = (1 + 0.3 * rng.standard_normal((ncoils, p.nk)) + 0.2j * rng.standard_normal((ncoils, p.nk)))
    call("imp_profile_enable", ctx.h, C.c_int(1))
    call("imp_profile_reset", ctx.h)
    got = ops.nudft_adjoint(ctx, kx, p.ky, p.kz, p.t, p.x, p.y, p.z, p.fm, d, p.sens, phi)
    tot, n = C.c_double(), C.c_int()
    call("imp_profile_read", ctx.h, b"k_nudft_stack", C.byref(tot), C.byref(n))
    call("imp_profile_enable", ctx.h, C.c_int(0))
    ref = orc.nudft_adjoint(kx, p.ky, p.kz, p.t, p.x, p.y, p.z, p.fm, d, p.sens, phi)
    err = rel_l2(got, ref)
    print(f"adjoint stack N={N} Nz={Nz} C={ncoils} perturbed={perturb}: "
          f"{n.value} stacked launches, rel_l2={err:.2e}")
    assert (n.value > 0) == (not perturb)
    assert err < OP_TOL


@pytest.mark.parametrize("N,Nz,ncoils,perturb", [(8, 8, 3, False), (8, 8, 3, True),
                                                 (12, 16, 4, False)])
def test_forward_stacked_trajectory_path(ctx, N, Nz, ncoils, perturb):
    """The forward NUDFT of a stack runs as a 2D NUDFT of one partition with
    the coil-partition pairs as right-hand sides (their weights carry
    e^{-i 2pi kz_p z}); a perturbed sample sends it to the 3D kernel."""
    import ctypes as C
    from impatient import ops
    from impatient._lib import call
    from oracle import oracle as orc
    p = _prob(N=N, Nz=Nz, ncoils=ncoils)
    t = p.t.copy()
    if perturb:
        t[len(t) // 2 + 5] += np.float32(1e-6)
    rng = np.random.default_rng(13)
    rho = (rng.standard_normal(p.nvox) + 1j * rng.standard_normal(p.nvox)).astype(np.complex64)
    call("imp_profile_enable", ctx.h, C.c_int(1))
    call("imp_profile_reset", ctx.h)
    got = ops.nudft_forward(ctx, p.kx, p.ky, p.kz, t, p.x, p.y, p.z, p.fm, rho, p.sens, ncoils)
    tot, n = C.c_double(), C.c_int()
    call("imp_profile_read", ctx.h, b"k_nudft_stack", C.byref(tot), C.byref(n))
    call("imp_profile_enable", ctx.h, C.c_int(0))
    ref = orc.nudft_forward(p.kx, p.ky, p.kz, t, p.x, p.y, p.z, p.fm, rho, p.sens)
    err = rel_l2(got, ref)
    print(f"forward stack N={N} Nz={Nz} C={ncoils} perturbed={perturb}: "
          f"{n.value} stacked launches, rel_l2={err:.2e}")
    assert (n.value > 0) == (not perturb)
    assert got.shape == (ncoils, p.nk)
    assert err < OP_TOL


@pytest.mark.parametrize("N,Nz,ncoils", [(16, 1, 1), (16, 1, 4), (8, 4, 3)])
def test_forward_matches_oracle(ctx, N, Nz, ncoils):
    from impatient import ops
    from oracle import oracle as orc
    p = _prob(N=N, Nz=Nz, ncoils=ncoils)
    rng = np.random.default_rng(2)
    rho = (rng.standard_normal(p.nvox) + 1j * rng.standard_normal(p.nvox)).astype(np.complex64)
    got = ops.nudft_forward(ctx, p.kx, p.ky, p.kz, p.t, p.x, p.y, p.z, p.fm, rho, p.sens, ncoils)
    ref = orc.nudft_forward(p.kx, p.ky, p.kz, p.t, p.x, p.y, p.z, p.fm, rho, p.sens)
    err = rel_l2(got, ref)
    print(f"forward N={N} Nz={Nz} C={ncoils}: rel_l2={err:.2e}")
    assert got.shape == (ncoils, p.nk)
    assert err < OP_TOL


def test_adjoint_identity_large(ctx):
    """Re<F a, b> == Re<a, F^H b> at a size the oracle does not run
    (64x64 image, 8 coils, 16384 samples)."""
    from impatient import ops, synth
    p = synth.make_problem("adj64", 64, 1, 8, 16, 1024, 8e-3, 150.0, 3, 4, 99)
    rng = np.random.default_rng(3)
    a = (rng.standard_normal(p.nvox) + 1j * rng.standard_normal(p.nvox)).astype(np.complex64)
    b = (rng.standard_normal((8, p.nk)) + 1j * rng.standard_normal((8, p.nk))).astype(np.complex64)
    Fa = ops.nudft_forward(ctx, p.kx, p.ky, None, p.t, p.x, p.y, None, p.fm, a, p.sens, 8)
    FHb = ops.nudft_adjoint(ctx, p.kx, p.ky, None, p.t, p.x, p.y, None, p.fm, b, p.sens)
    lhs = np.vdot(b.astype(np.complex128), Fa.astype(np.complex128))
    rhs = np.vdot(FHb.astype(np.complex128), a.astype(np.complex128))
    print(f"<b,Fa>={lhs:.6e} <F^H b,a>={rhs:.6e}")
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


def test_adjoint_linearity_and_coil_sum(ctx):
    """F^H(2 d1 + d2) = 2 F^H d1 + F^H d2; with S = 1 the C-coil adjoint of
    identical coils is C times the single-coil adjoint."""
    from impatient import ops
    p = _prob(N=16, ncoils=1)
    rng = np.random.default_rng(4)
    d1 = (rng.standard_normal(p.nk) + 1j * rng.standard_normal(p.nk)).astype(np.complex64)
    d2 = (rng.standard_normal(p.nk) + 1j * rng.standard_normal(p.nk)).astype(np.complex64)
    f = lambda d: ops.nudft_adjoint(ctx, p.kx, p.ky, None, p.t, p.x, p.y, None, p.fm, d)
    assert rel_l2(f((2 * d1 + d2).astype(np.complex64)), 2 * f(d1) + f(d2)) < 1e-5
    d3 = np.stack([d1, d1, d1])
    assert rel_l2(f(d3), 3 * f(d1)) < 1e-6


@pytest.mark.parametrize("N,Nz,L", [(8, 1, 0), (8, 1, 3), (8, 1, 12), (4, 4, 2)])
def test_q_matches_oracle(ctx, N, Nz, L):
    from impatient import ops
    from oracle import oracle as orc
    p = _prob(N=N, Nz=Nz, ncoils=1, ntime_segs=L)
    got = ops.toeplitz_q(ctx, p.kx, p.ky, p.kz, p.t, p.Nx, p.Ny, p.Nz, L)
    ref = orc.toeplitz_q(p.kx, p.ky, p.kz, p.t, p.Nx, p.Ny, p.Nz, L)
    err = rel_l2(got, ref)
    print(f"Q N={N} Nz={Nz} L={L}: rel_l2={err:.2e}")
    assert got.shape == ref.shape
    assert err < OP_TOL


@pytest.mark.parametrize("N,Nz,L,phi_mode", [(8, 4, 3, None), (12, 8, 4, "stack"),
                                              (8, 4, 3, "free")])
def test_q_stacked_trajectory_path(ctx, N, Nz, L, phi_mode):
    """Q_l of a stack is the 2D kernel of one partition times
    Z(z) = sum_p e^{+i 2pi kz_p z} (k_q_outer).  Density weights that repeat
    per partition keep the stacked path; weights that differ between
    partitions send it to the 3D kernels.  All match the FP64 oracle."""
    import ctypes as C
    from impatient import ops
    from impatient._lib import call
    from oracle import oracle as orc
    p = _prob(N=N, Nz=Nz, ncoils=1, ntime_segs=L)
    rng = np.random.default_rng(12)
    K = p.nk // Nz
    phi = None
    if phi_mode == "stack":
        phi = np.tile(1 + 0.3 * rng.standard_normal(K) + 0.2j * rng.standard_normal(K), Nz)
    elif phi_mode == "free":
        phi = 1 + 0.3 * rng.standard_normal(p.nk) + 0.2j * rng.standard_normal(p.nk)
    call("imp_profile_enable", ctx.h, C.c_int(1))
    call("imp_profile_reset", ctx.h)
    got = ops.toeplitz_q(ctx, p.kx, p.ky, p.kz, p.t, N, N, Nz, L, phi=phi)
    tot, n = C.c_double(), C.c_int()
    call("imp_profile_read", ctx.h, b"k_q_outer", C.byref(tot), C.byref(n))
    call("imp_profile_enable", ctx.h, C.c_int(0))
    ref = orc.toeplitz_q(p.kx, p.ky, p.kz, p.t, N, N, Nz, L, phi=phi)
    err = rel_l2(got, ref)
    print(f"Q stack N={N} Nz={Nz} L={L} phi={phi_mode}: {n.value} outer-product launches, "
          f"rel_l2={err:.2e}")
    assert (n.value == 1) == (phi_mode != "free")
    assert err < OP_TOL


@pytest.mark.parametrize("Nx,Ny,Nz", [(6, 10, 1), (5, 3, 3), (4, 1, 2)])
def test_q_nonsquare_grid_and_hermitian_fill(ctx, Nx, Ny, Nz):
    """The standard grid is evaluated on half the points and mirrored:
    check every point against the FP64 oracle on non-square grids (mirror
    edges x = -Nx, z = -Nz; Ny = 1 has no mirrored row) and that every
    point with a mirror on the grid, the row y = 0 included, is the exact
    conjugate of it (Q(0) real)."""
    from impatient import ops
    from oracle import oracle as orc
    rng = np.random.default_rng(5)
    M = 300
    kx = rng.uniform(-Nx / 2, Nx / 2, M).astype(np.float32)
    ky = rng.uniform(-Ny / 2, Ny / 2, M).astype(np.float32)
    kz = rng.uniform(-Nz / 2, Nz / 2, M).astype(np.float32) if Nz > 1 else None
    t = np.linspace(0, 4e-3, M, dtype=np.float32)
    L = 3
    got = ops.toeplitz_q(ctx, kx, ky, kz, t, Nx, Ny, Nz, L)
    ref = orc.toeplitz_q(kx, ky, kz, t, Nx, Ny, Nz, L)
    err = rel_l2(got, ref)
    print(f"Q {Nx}x{Ny}x{Nz}: rel_l2={err:.2e}")
    assert err < OP_TOL
    # every point whose mirror is on the grid (y > -Ny, x > -Nx, z > -Nz)
    M3 = 2 * Nz if Nz > 1 else 1
    g = got.reshape(L + 1, 2 * Ny, 2 * Nx, M3)
    z0 = 1 if M3 > 1 else 0
    for iy in range(1, 2 * Ny):
        for ix in range(1, 2 * Nx):
            for iz in range(z0, M3):
                mz = 2 * Nz - iz if M3 > 1 else 0
                assert np.array_equal(g[:, iy, ix, iz], np.conj(g[:, 2 * Ny - iy, 2 * Nx - ix, mz]))
    assert np.all(g[:, Ny, Nx, Nz if M3 > 1 else 0].imag == 0)


def test_q_cumulative_compat(ctx):
    """IMP_Q_CUMULATIVE returns the prefix sums over l (reference direct path)."""
    from impatient import ops, IMP_Q_CUMULATIVE
    p = _prob(N=8, ncoils=1, ntime_segs=4)
    q = ops.toeplitz_q(ctx, p.kx, p.ky, None, p.t, 8, 8, 1, 4)
    qc = ops.toeplitz_q(ctx, p.kx, p.ky, None, p.t, 8, 8, 1, 4, flags=IMP_Q_CUMULATIVE)
    assert rel_l2(qc, np.cumsum(q.astype(np.complex128), axis=0)) < 1e-6


def test_q_explicit_grid_matches_standard(ctx):
    from impatient import ops
    from oracle import oracle as orc
    p = _prob(N=8, ncoils=1, ntime_segs=2)
    g = orc.q_grid(8, 8, 1)
    a = ops.toeplitz_q(ctx, p.kx, p.ky, None, p.t, 8, 8, 1, 2)
    b = ops.toeplitz_q(ctx, p.kx, p.ky, None, p.t, 8, 8, 1, 2, grid=(g[0], g[1], None))
    assert rel_l2(a, b) < 1e-6


def test_adjoint_batch_matches_single_calls_and_oracle(ctx):
    """Three images of different sizes, trajectories and coil counts in one
    persistent launch: each equals its own imp_nudft_adjoint call (same
    arithmetic, different chunking) and the FP64 oracle."""
    from impatient import ops, synth
    from oracle import oracle as orc
    probs = [synth.tiny(N=16, ncoils=4, seed=1), synth.tiny(N=24, ncoils=6, seed=2, J=6),
             synth.tiny(N=32, ncoils=8, seed=3, J=8, S=256)]
    jobs = []
    for p in probs:
        y = orc.nudft_forward(p.kx, p.ky, None, p.t, p.x, p.y, None, p.fm, p.rho, p.sens)
        jobs.append(dict(kx=p.kx, ky=p.ky, t=p.t, x=p.x, y=p.y, fm=p.fm, d=y, sens=p.sens))
    for scale in (0, 2):
        got = ops.nudft_adjoint_batch(ctx, jobs, chunk_scale=scale)
        for p, jb, g in zip(probs, jobs, got):
            one = ops.nudft_adjoint(ctx, p.kx, p.ky, None, p.t, p.x, p.y, None, p.fm, jb["d"],
                                    p.sens)
            ref = orc.nudft_adjoint(p.kx, p.ky, None, p.t, p.x, p.y, None, p.fm, jb["d"], p.sens)
            e1, e2 = rel_l2(g, one), rel_l2(g, ref)
            print(f"batch N={p.Nx} C={p.ncoils} scale={scale}: vs single {e1:.2e}, vs FP64 {e2:.2e}")
            assert e1 < 1e-6 and e2 < OP_TOL


@pytest.mark.parametrize("Nx,Ny,Nz,L", [(8, 8, 1, 3), (5, 3, 1, 0), (12, 12, 1, 12), (4, 4, 2, 2),
                                        (3, 5, 3, 1)])
def test_q_separable_and_direct_paths_match_oracle(ctx, Nx, Ny, Nz, L):
    """The standard grid's separable GEMM path (default) and the per-point
    path (IMP_Q_DIRECT) against the FP64 oracle, with |phi|^2 weights and a
    readout whose times are not monotone and partly outside [t0, t_last]:
    samples before t0 - tau or after t0 + (L+1) tau have zero weight in every
    segment and samples in (t0 - tau, t0) feed segment 0 only."""
    from impatient import ops, IMP_Q_DIRECT
    from oracle import oracle as orc
    rng = np.random.default_rng(40 + Nx + L)
    M = 700
    kx = rng.uniform(-Nx / 2, Nx / 2, M).astype(np.float32)
    ky = rng.uniform(-Ny / 2, Ny / 2, M).astype(np.float32)
    kz = rng.uniform(-Nz / 2, Nz / 2, M).astype(np.float32) if Nz > 1 else None
    t = rng.uniform(-2e-3, 6e-3, M).astype(np.float32)
    t[0], t[-1] = 0.0, 4e-3
    phi = (1 + 0.3 * rng.standard_normal(M)) + 0.2j * rng.standard_normal(M)
    ref = orc.toeplitz_q(kx, ky, kz, t, Nx, Ny, Nz, L, phi=phi)
    for flags in (0, IMP_Q_DIRECT):
        got = ops.toeplitz_q(ctx, kx, ky, kz, t, Nx, Ny, Nz, L, phi=phi, flags=flags)
        err = rel_l2(got, ref)
        print(f"Q {Nx}x{Ny}x{Nz} L={L} flags={flags}: rel_l2={err:.2e}")
        assert err < 2e-6


def test_q_separable_matches_direct_at_larger_sizes(ctx):
    """GEMM path against the per-point path where the FP64 oracle is too slow:
    a 128x128 2D grid (rows span two workgroups of four waves, columns 16
    tiles) with L = 10 and a 3D 24x16x6 grid (rows not a multiple of 64)."""
    from impatient import ops, synth, IMP_Q_DIRECT
    cases = [(synth.make_problem("q128", 128, 1, 1, 8, 2048, 8e-3, 100.0, 3, 10, 5), 10),
             (synth.make_problem("q3d", 16, 6, 1, 4, 512, 5e-3, 100.0, 2, 4, 6), 4)]
    for p, L in cases:
        kz = p.kz if p.Nz > 1 else None
        # non-square 3D grid: stretch x
        Nx = p.Nx if p.Nz == 1 else 24
        a = ops.toeplitz_q(ctx, p.kx, p.ky, kz, p.t, Nx, p.Ny, p.Nz, L)
        b = ops.toeplitz_q(ctx, p.kx, p.ky, kz, p.t, Nx, p.Ny, p.Nz, L, flags=IMP_Q_DIRECT)
        err = rel_l2(a, b)
        print(f"Q {Nx}x{p.Ny}x{p.Nz} L={L}: separable vs direct rel_l2={err:.2e}")
        assert err < 2e-6


def test_q_separable_cumulative_matches_direct(ctx):
    from impatient import ops, IMP_Q_CUMULATIVE, IMP_Q_DIRECT
    p = _prob(N=16, ncoils=1, ntime_segs=5)
    a = ops.toeplitz_q(ctx, p.kx, p.ky, None, p.t, 16, 16, 1, 5, flags=IMP_Q_CUMULATIVE)
    b = ops.toeplitz_q(ctx, p.kx, p.ky, None, p.t, 16, 16, 1, 5,
                       flags=IMP_Q_CUMULATIVE | IMP_Q_DIRECT)
    assert rel_l2(a, b) < 2e-6


@pytest.mark.parametrize("N", [8, 16])
def test_q_separable_split_is_consistent_per_segment(ctx, N):
    """Every segment of the GEMM path within 1e-6 of the per-point path.  With
    L = 12 these readouts contain products w * ey whose FP32 rounding and
    direct fp16 rounding differ (segment 6 row y = -1/8 at N = 8, segment 10
    row y = -1/16 at N = 16): the fp16 hi/lo split must take both parts from
    the same rounded product, else that sample is off by one fp16 ulp
    (~1e-5 of the segment)."""
    from impatient import ops, synth, IMP_Q_DIRECT
    p = synth.tiny(N=N, ncoils=1, ntime_segs=12)
    a = ops.toeplitz_q(ctx, p.kx, p.ky, None, p.t, N, N, 1, 12).reshape(13, -1)
    b = ops.toeplitz_q(ctx, p.kx, p.ky, None, p.t, N, N, 1, 12, flags=IMP_Q_DIRECT).reshape(13, -1)
    per = [rel_l2(a[l], b[l]) for l in range(13)]
    print(f"N={N} per-segment separable vs direct: " + " ".join(f"{e:.1e}" for e in per))
    assert max(per) < 1e-6
