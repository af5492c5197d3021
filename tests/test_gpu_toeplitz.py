"""GPU parity of the Toeplitz normal operator, the FD regularizer and CG."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _setup(N=16, Nz=1, ncoils=3, L=3, seed=7, max_hz=60.0):
    from impatient import synth
    from oracle import oracle as orc
    p = synth.tiny(N=N, Nz=Nz, ncoils=ncoils, ntime_segs=L, seed=seed, max_hz=max_hz)
    q = orc.toeplitz_q(p.kx, p.ky, p.kz, p.t, p.Nx, p.Ny, p.Nz, L)
    tau, t0 = orc.segment_timing(p.t, L)
    return p, q, tau, t0


@pytest.mark.parametrize("N,Nz,ncoils,L,lam", [(16, 1, 1, 0, 0.0), (16, 1, 3, 3, 1e-2),
                                                (32, 1, 2, 5, 0.0), (8, 4, 2, 2, 1e-2),
                                                (24, 1, 2, 3, 1e-2), (48, 1, 1, 2, 0.0),
                                                (8, 24, 2, 2, 1e-2)])
def test_apply_matches_oracle(ctx, N, Nz, ncoils, L, lam):
    from impatient import ops
    from oracle import oracle as orc
    p, q, tau, t0 = _setup(N, Nz, ncoils, L)
    shape = p.shape if Nz > 1 else p.shape[:2]
    op = ops.Toeplitz(ctx, p.shape, q.astype(np.complex64), tau, t0, p.fm, p.sens, lam)
    rng = np.random.default_rng(5)
    v = (rng.standard_normal(p.nvox) + 1j * rng.standard_normal(p.nvox)).astype(np.complex64)
    got = op.apply(v)
    ref = orc.toeplitz_apply(v.reshape(shape), q, tau, t0, p.fm, p.sens, lam).reshape(-1)
    err = rel_l2(got, ref)
    print(f"toeplitz apply N={N} Nz={Nz} C={ncoils} L={L}: rel_l2={err:.2e}")
    assert err < 2e-6


@pytest.mark.parametrize("N,Nz", [(16, 1), (8, 4)])
def test_apply_real_and_complex_spectrum_paths(ctx, N, Nz):
    """A Q_l that is exactly conjugate-symmetric (the GPU's standard-grid
    Q_l) has a real spectrum once the offsets -N, which never reach the
    cropped product, are zeroed: the apply then multiplies by Re(Qf_l).  Any
    other Q_l keeps the complex spectrum.  Both match the oracle's apply on
    the same Q_l: the GPU's own Q_l, and that Q_l plus a 10% perturbation
    that is not conjugate-symmetric (with the real path taken by mistake it
    would be ~1e-1 off)."""
    from impatient import ops
    from oracle import oracle as orc
    p, _, tau, t0 = _setup(N, Nz, 2, 3)
    q = ops.toeplitz_q(ctx, p.kx, p.ky, p.kz, p.t, p.Nx, p.Ny, p.Nz, 3)
    rng = np.random.default_rng(9)
    pert = rng.standard_normal(q.shape) + 1j * rng.standard_normal(q.shape)
    q_asym = (q + 0.1 * np.abs(q).max() * pert).astype(np.complex64)
    shape = p.shape if Nz > 1 else p.shape[:2]
    v = (rng.standard_normal(p.nvox) + 1j * rng.standard_normal(p.nvox)).astype(np.complex64)
    for name, qq in (("conjugate-symmetric", q), ("perturbed", q_asym)):
        op = ops.Toeplitz(ctx, p.shape, qq, tau, t0, p.fm, p.sens, 1e-2)
        got = op.apply(v)
        op.close()
        ref = orc.toeplitz_apply(v.reshape(shape), qq.astype(np.complex128), tau, t0, p.fm,
                                 p.sens, 1e-2).reshape(-1)
        err = rel_l2(got, ref)
        print(f"N={N} Nz={Nz} {name} Q_l: apply vs oracle {err:.2e}")
        assert err < 2e-6


def test_apply_is_normal_operator_without_fieldmap(ctx):
    """With w = 0 and L = 1 (partition of unity), the Toeplitz apply is
    exactly F^H F: compare with forward then adjoint on the GPU."""
    from impatient import ops, synth
    from oracle import oracle as orc
    p = synth.tiny(N=16, ncoils=2, ntime_segs=1, max_hz=0.0)
    fm0 = np.zeros_like(p.fm)
    q = ops.toeplitz_q(ctx, p.kx, p.ky, None, p.t, 16, 16, 1, 1)
    tau, t0 = orc.segment_timing(p.t, 1)
    op = ops.Toeplitz(ctx, p.shape, q, tau, t0, fm0, p.sens, 0.0)
    rng = np.random.default_rng(6)
    v = (rng.standard_normal(p.nvox) + 1j * rng.standard_normal(p.nvox)).astype(np.complex64)
    Fv = ops.nudft_forward(ctx, p.kx, p.ky, None, p.t, p.x, p.y, None, fm0, v, p.sens, 2)
    FHFv = ops.nudft_adjoint(ctx, p.kx, p.ky, None, p.t, p.x, p.y, None, fm0, Fv, p.sens)
    err = rel_l2(op.apply(v), FHFv)
    print(f"T vs F^H F: {err:.2e}")
    assert err < 1e-5


@pytest.mark.parametrize("cfg,Ls", [(0, (2, 4, 8, 12)), (1, (2, 4, 6, 12))])
def test_time_segmentation_model_error(ctx, cfg, Ls):
    """Model error of the time-segmented Toeplitz operator (the reference's
    approximation of F^H F with a field map, L + 1 Hanning interpolators)
    against the exact F^H F (forward then adjoint NUDFT with the field map),
    on the configuration's phantom, trajectory, coils and field map: stated
    per L; it falls with L (configuration 0 runs at the reference default
    L = 12, configuration 1 at L = 6)."""
    from impatient import ops, synth
    from oracle import oracle as orc
    p = synth.config(cfg)
    v = p.rho.astype(np.complex64)
    nc = p.ncoils
    Fv = ops.nudft_forward(ctx, p.kx, p.ky, p.kz, p.t, p.x, p.y, p.z, p.fm, v, p.sens, nc)
    exact = ops.nudft_adjoint(ctx, p.kx, p.ky, p.kz, p.t, p.x, p.y, p.z, p.fm, Fv, p.sens)
    errs = {}
    for L in Ls:
        tau, t0 = orc.segment_timing(p.t, L)
        q = ops.toeplitz_q(ctx, p.kx, p.ky, p.kz, p.t, p.Nx, p.Ny, p.Nz, L)
        op = ops.Toeplitz(ctx, p.shape, q, tau, t0, p.fm, p.sens, 0.0)
        errs[L] = rel_l2(op.apply(v), exact)
        op.close()
    print(f"{p.name} time-segmentation model error: "
          + ", ".join(f"L={L} {e:.2e}" for L, e in errs.items()))
    vals = [errs[L] for L in Ls]
    assert all(b < a for a, b in zip(vals, vals[1:]))
    assert errs[Ls[-1]] < 1e-2


@pytest.mark.parametrize("shape,weighted", [((16, 16), False), ((16, 16), True),
                                            ((8, 12, 4), False), ((8, 8, 6), True)])
def test_fd_reg_matches_oracle(ctx, shape, weighted):
    from impatient import ops
    from oracle import oracle as orc
    rng = np.random.default_rng(8)
    v = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    w = rng.uniform(0.1, 2.0, shape).astype(np.float32) if weighted else None
    got = ops.fd_reg(ctx, v, 0.37, w)
    ref = orc.fd_reg(v, 0.37, w)
    assert rel_l2(got, ref) < 1e-6


def test_tv_weight_matches_oracle(ctx):
    from impatient import ops
    from oracle import oracle as orc
    rng = np.random.default_rng(9)
    for shape in [(16, 16), (8, 8, 4)]:
        v = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
        assert rel_l2(ops.tv_weight(ctx, v), orc.tv_weight(v)) < 1e-6


def test_cg_matches_oracle_iterates(ctx):
    """10 CG iterations against the oracle CG of the same FP64 operator
    (toeplitz_recon semantics; b = F^H y, 4 coils, 8192 samples, lambda = 1),
    best and last iterate.  The reference CG keeps its vectors in FP32, and
    an FP32-vector run of the oracle already ends 3.1e-2 from the all-FP64
    one here, so the GPU is pinned to the FP32-vector oracle run (measured
    9.2e-5) and its distance to the FP64 run is bounded by that run's own
    distance plus twice the pin (triangle inequality, with margin)."""
    from impatient import ops, synth, IMP_CG_BEST_ITERATE
    from oracle import oracle as orc
    p = synth.tiny(N=16, ncoils=4, ntime_segs=3, J=16, S=512)
    q = orc.toeplitz_q(p.kx, p.ky, None, p.t, 16, 16, 1, 3)
    tau, t0 = orc.segment_timing(p.t, 3)
    lam = 1.0
    y = orc.nudft_forward(p.kx, p.ky, None, p.t, p.x, p.y, None, p.fm, p.rho, p.sens)
    b = orc.nudft_adjoint(p.kx, p.ky, None, p.t, p.x, p.y, None, p.fm, y, p.sens)
    op = ops.Toeplitz(ctx, p.shape, q.astype(np.complex64), tau, t0, p.fm, p.sens, lam)
    A = lambda v: orc.toeplitz_apply(v.reshape(16, 16), q, tau, t0, p.fm, p.sens, lam).reshape(-1)
    for best in (True, False):
        x, st = op.cg(b.astype(np.complex64), 10, 0.0, IMP_CG_BEST_ITERATE if best else 0)
        x32, it, hist32 = orc.toeplitz_cg(A, b, 10, best_iterate=best, vec_dtype=np.complex64)
        x64, it, hist64 = orc.toeplitz_cg(A, b, 10, best_iterate=best)
        e32, e64 = rel_l2(x, x32), rel_l2(x, x64)
        print(f"CG 10 iters best={best}: vs FP32-storage CG {e32:.2e}, vs FP64 CG {e64:.2e}, "
              f"res {st['rel_residual']:.3e} vs {hist32[-1]:.3e} / {hist64[-1]:.3e}")
        spread = rel_l2(x32, x64)
        assert st["iterations"] == 10
        assert abs(st["rel_residual"] - hist32[-1]) / hist32[-1] < 5e-3
        assert e32 < 1e-3
        assert e64 <= spread + 2 * e32


def test_cg_converges_to_dense_solution(ctx):
    """CG to tol 1e-6 on b = F^H y matches the dense solve of the FP64
    operator (8x8 image, 2 coils, 2048 samples, lambda 0.05)."""
    from impatient import ops, synth, IMP_CG_BEST_ITERATE, IMP_CG_TRUE_RESIDUAL
    from oracle import oracle as orc
    p = synth.tiny(N=8, ncoils=2, ntime_segs=2, J=8, S=256)
    q = orc.toeplitz_q(p.kx, p.ky, None, p.t, 8, 8, 1, 2)
    tau, t0 = orc.segment_timing(p.t, 2)
    lam = 5e-2
    op = ops.Toeplitz(ctx, p.shape, q.astype(np.complex64), tau, t0, p.fm, p.sens, lam)
    n = p.nvox
    A = np.zeros((n, n), dtype=np.complex128)
    for j in range(n):
        e = np.zeros(n, np.complex128)
        e[j] = 1
        A[:, j] = orc.toeplitz_apply(e.reshape(8, 8), q, tau, t0, p.fm, p.sens, lam).reshape(-1)
    y = orc.nudft_forward(p.kx, p.ky, None, p.t, p.x, p.y, None, p.fm, p.rho, p.sens)
    b = orc.nudft_adjoint(p.kx, p.ky, None, p.t, p.x, p.y, None, p.fm, y, p.sens)
    xd = np.linalg.solve(A, b)
    x, st = op.cg(b.astype(np.complex64), 500, 1e-6, IMP_CG_BEST_ITERATE | IMP_CG_TRUE_RESIDUAL)
    err = rel_l2(x, xd)
    # forward-error bound: |x - xd| / |xd| <= cond(A) * |A x - b| / |b|
    # with the FP64 operator (so FP32 operator error is inside the residual)
    r64 = np.linalg.norm(A @ x.astype(np.complex128) - b) / np.linalg.norm(b)
    kappa = np.linalg.cond(A)
    print(f"CG dense: iters={st['iterations']} true_res={st['true_rel_residual']:.2e} "
          f"FP64 residual {r64:.2e} cond {kappa:.1f} image vs dense FP64 {err:.2e}")
    assert st["converged"] == 1
    assert st["true_rel_residual"] < 1e-5
    assert r64 < 2e-6
    assert err <= kappa * r64 * 1.01


def test_cg_continues_through_negative_curvature_like_the_reference(ctx):
    """An indefinite Hermitian operator (lambda = -1000 makes the regularizer
    dominate with the wrong sign): p^H A p < 0 at the first and fourth
    iteration.  The CG takes alpha = rho / p^H A p of either sign as the
    reference loop does (recon.cu:2071-2077) and matches the oracle CG
    iterate by iterate instead of stopping."""
    from impatient import ops, synth
    from oracle import oracle as orc
    p = synth.tiny(N=8, ncoils=2, ntime_segs=2, J=8, S=256)
    q = orc.toeplitz_q(p.kx, p.ky, None, p.t, 8, 8, 1, 2)
    tau, t0 = orc.segment_timing(p.t, 2)
    lam = -1000.0
    op = ops.Toeplitz(ctx, p.shape, q.astype(np.complex64), tau, t0, p.fm, p.sens, lam)
    A = lambda v: orc.toeplitz_apply(v.reshape(8, 8), q, tau, t0, p.fm, p.sens, lam).reshape(-1)
    rng = np.random.default_rng(12)
    b = rng.standard_normal(p.nvox) + 1j * rng.standard_normal(p.nvox)
    assert np.vdot(b, A(b)).real < 0   # first search direction p = b
    x, st = op.cg(b.astype(np.complex64), 4, 0.0, 0)
    x32, _, h32 = orc.toeplitz_cg(A, b, 4, best_iterate=False, vec_dtype=np.complex64)
    x64, _, h64 = orc.toeplitz_cg(A, b, 4, best_iterate=False)
    e32, e64 = rel_l2(x, x32), rel_l2(x, x64)
    print(f"indefinite CG 4 iters: vs FP32-storage CG {e32:.2e}, vs FP64 CG {e64:.2e}, "
          f"res {st['rel_residual']:.3e} vs {h32[-1]:.3e} / {h64[-1]:.3e}")
    assert st["iterations"] == 4
    assert abs(st["rel_residual"] - h32[-1]) / h32[-1] < 1e-4
    assert e32 < 1e-5 and e64 < 1e-5
