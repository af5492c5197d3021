"""GPU parity of the sparse-matrix regularizer lambda C^T W C (the -reg
option): the CSR kernels against the oracle, C = the periodic difference
matrix against the finite-difference stencil, and both solvers (Toeplitz CG,
brute-force CG) with C in place of the stencil."""
import numpy as np
import pytest

from conftest import difference_coo, rel_l2

pytestmark = pytest.mark.gpu


def random_coo(rows, cols, per_row, seed):
    rng = np.random.default_rng(seed)
    I = np.repeat(np.arange(rows), per_row)
    J = rng.integers(0, cols, I.size)
    V = rng.standard_normal(I.size)
    I = np.concatenate([I, I[:7]])          # duplicates add
    J = np.concatenate([J, J[:7]])
    V = np.concatenate([V, V[:7]])
    return rows, cols, I.astype(np.int32), J.astype(np.int32), V


@pytest.mark.parametrize("shape,weighted", [((16, 12), False), ((16, 12), True),
                                            ((8, 6, 5), True)])
def test_apply_matches_oracle(ctx, shape, weighted):
    from impatient import ops
    from oracle import oracle as orc
    n = int(np.prod(shape))
    coo = random_coo(len(shape) * n, n, 3, seed=11)
    rng = np.random.default_rng(12)
    p = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    w = rng.uniform(0.2, 2.0, shape).astype(np.float32) if weighted else None
    R = ops.SparseReg(ctx, *coo)
    got = R.apply(p, 0.37, w)
    ref = orc.sparse_reg(p, 0.37, *coo, w)
    err = rel_l2(got, ref)
    print(f"sparse reg {shape} weighted={weighted}: {err:.2e}")
    assert err < 1e-6
    R.close()


@pytest.mark.parametrize("shape", [(16, 12), (8, 6, 5)])
def test_difference_matrix_equals_stencil(ctx, shape):
    from impatient import ops
    from oracle import oracle as orc
    rng = np.random.default_rng(13)
    p = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    w = rng.uniform(0.2, 2.0, shape).astype(np.float32)
    R = ops.SparseReg(ctx, *difference_coo(shape))
    got = R.apply(p, 0.5, w)
    stencil = ops.fd_reg(ctx, p, 0.5, w)
    e_st, e_or = rel_l2(got, stencil), rel_l2(got, orc.fd_reg(p, 0.5, w))
    print(f"C = D {shape}: vs stencil {e_st:.2e}, vs oracle {e_or:.2e}")
    assert e_st < 1e-6 and e_or < 1e-6


def test_toeplitz_operator_with_sparse_reg(ctx):
    """A p = sum_c S_c^H T S_c p + lambda C^T C p for a random C; None
    restores the stencil; a C of the wrong shape is rejected."""
    from impatient import _lib, ops, synth
    from oracle import oracle as orc
    L, lam = 3, 2e-2
    p = synth.tiny(N=16, ncoils=3, ntime_segs=L, seed=7, max_hz=60.0)
    q = orc.toeplitz_q(p.kx, p.ky, p.kz, p.t, p.Nx, p.Ny, p.Nz, L)
    tau, t0 = orc.segment_timing(p.t, L)
    shape = p.shape[:2]
    op = ops.Toeplitz(ctx, p.shape, q.astype(np.complex64), tau, t0, p.fm, p.sens, lam)
    rng = np.random.default_rng(14)
    v = (rng.standard_normal(p.nvox) + 1j * rng.standard_normal(p.nvox)).astype(np.complex64)
    coo = random_coo(2 * p.nvox, p.nvox, 2, seed=15)
    R = ops.SparseReg(ctx, *coo)
    op.set_sparse_reg(R)
    got = op.apply(v)
    ref = (orc.toeplitz_apply(v.reshape(shape), q, tau, t0, p.fm, p.sens, 0.0).reshape(-1)
           + orc.sparse_reg(v, lam, *coo))
    err = rel_l2(got, ref)
    op.set_sparse_reg(None)
    err_fd = rel_l2(op.apply(v), orc.toeplitz_apply(v.reshape(shape), q, tau, t0, p.fm,
                                                     p.sens, lam).reshape(-1))
    print(f"toeplitz + sparse reg: {err:.2e}; stencil restored: {err_fd:.2e}")
    assert err < 2e-6 and err_fd < 2e-6
    bad = ops.SparseReg(ctx, *random_coo(p.nvox, p.nvox, 2, seed=16))
    with pytest.raises(_lib.ImpError) as e:
        op.set_sparse_reg(bad)
    assert e.value.status == -1
    # CG with C = D follows CG with the stencil
    D = ops.SparseReg(ctx, *difference_coo(shape))
    b = orc.toeplitz_apply(p.rho.reshape(shape), q, tau, t0, p.fm, p.sens,
                           lam).reshape(-1).astype(np.complex64)
    x_fd, st_fd = op.cg(b, 6)
    op.set_sparse_reg(D)
    x_d, st_d = op.cg(b, 6)
    e_cg = rel_l2(x_d, x_fd)
    print(f"CG(6) C = D vs stencil: {e_cg:.2e}")
    assert e_cg < 1e-4
    # the graph-captured iteration replays the same kernels (random C, twice)
    from impatient._lib import IMP_CG_BEST_ITERATE, IMP_CG_GRAPH
    op.set_sparse_reg(R)
    x_e, _ = op.cg(b, 6)
    for _ in range(2):
        x_g, _ = op.cg(b, 6, 0.0, IMP_CG_BEST_ITERATE | IMP_CG_GRAPH)
        e_g = rel_l2(x_g, x_e)
        print(f"CG(6) with C, graph vs eager: {e_g:.2e}")
        assert e_g < 1e-6
    # destroying the regularizer while the operator uses it: applies fail
    # (no dangling use) until the regularizer is replaced
    R.close()
    with pytest.raises(_lib.ImpError) as e2:
        op.apply(v)
    assert e2.value.status == -1
    op.set_sparse_reg(None)
    assert rel_l2(op.apply(v), orc.toeplitz_apply(v.reshape(shape), q, tau, t0, p.fm, p.sens,
                                                  lam).reshape(-1)) < 2e-6
    op.close()


def test_bruteforce_with_sparse_reg(ctx):
    """bruteForceGpu's -reg branch: R = lambda C^T C from C."""
    from impatient import ops, synth
    from oracle import oracle as orc
    lam, cg_num = 0.1, 8
    p = synth.tiny(N=16, ncoils=4, ntime_segs=3, J=16, S=512)
    y = orc.nudft_forward(p.kx, p.ky, None, p.t, p.x, p.y, None, p.fm, p.rho, p.sens)
    y = synth.add_noise(y, 40.0, 3)
    x0 = np.zeros(p.nvox, np.complex64)
    coo = random_coo(2 * p.nvox, p.nvox, 2, seed=17)
    R = ops.SparseReg(ctx, *coo)
    got = ops.bruteforce_cg(ctx, p.kx, p.ky, p.t, p.x, p.y, p.fm, y, p.sens, cg_num, lam, x0,
                            reg=R)
    F = lambda v: orc.nudft_forward(p.kx, p.ky, None, p.t, p.x, p.y, None, p.fm, v, p.sens)
    FH = lambda d: orc.nudft_adjoint(p.kx, p.ky, None, p.t, p.x, p.y, None, p.fm, d, p.sens)
    Rf = lambda v: orc.sparse_reg(v, lam, *coo)
    ref32 = orc.bruteforce_cg(F, FH, Rf, y, x0, cg_num, vec_dtype=np.complex64)
    err = rel_l2(got, ref32)
    D = ops.SparseReg(ctx, *difference_coo((16, 16)))
    x_d = ops.bruteforce_cg(ctx, p.kx, p.ky, p.t, p.x, p.y, p.fm, y, p.sens, cg_num, lam, x0,
                            reg=D)
    x_fd = ops.bruteforce_cg(ctx, p.kx, p.ky, p.t, p.x, p.y, p.fm, y, p.sens, cg_num, lam, x0)
    e_d = rel_l2(x_d, x_fd)
    print(f"brute force + sparse reg vs FP32-storage oracle {err:.2e}; C = D vs stencil {e_d:.2e}")
    assert err < 5e-3 and e_d < 1e-4


def test_regularizer_outlives_its_context(ctx):
    """Destroying the context first detaches the regularizer: its device
    memory is freed with the context, later calls report an error, and the
    final destroy frees only the host object."""
    from impatient import _lib, ops
    from impatient._lib import Context, ptr
    c2 = Context(0)
    coo = random_coo(32, 16, 2, seed=18)
    R = ops.SparseReg(c2, *coo)
    p = np.ones(16, np.complex64)
    ref = R.apply(p, 1.0)
    assert np.isfinite(ref).all()
    c2.close()
    a = [ctx.alloc((16,)) for _ in range(4)]   # valid arrays on a live context
    with pytest.raises(_lib.ImpError) as e:
        _lib.call("imp_sparse_reg_apply", R.h, _lib.C.c_float(1.0), None, ptr(a[0]),
                  ptr(a[1]), ptr(a[2]), ptr(a[3]))
    assert e.value.status == -1
    R.close()


def test_toeplitz_operator_with_sparse_reg_3d(ctx):
    """3D image: C has three blocks of N rows (recon.cu:1613-1631)."""
    from impatient import ops, synth
    from oracle import oracle as orc
    L, lam = 2, 1e-2
    p = synth.tiny(N=8, Nz=4, ncoils=2, ntime_segs=L, seed=7, max_hz=60.0)
    q = orc.toeplitz_q(p.kx, p.ky, p.kz, p.t, p.Nx, p.Ny, p.Nz, L)
    tau, t0 = orc.segment_timing(p.t, L)
    op = ops.Toeplitz(ctx, p.shape, q.astype(np.complex64), tau, t0, p.fm, p.sens, lam)
    rng = np.random.default_rng(19)
    v = (rng.standard_normal(p.nvox) + 1j * rng.standard_normal(p.nvox)).astype(np.complex64)
    coo = random_coo(3 * p.nvox, p.nvox, 2, seed=20)
    R = ops.SparseReg(ctx, *coo)
    op.set_sparse_reg(R)
    got = op.apply(v)
    ref = (orc.toeplitz_apply(v.reshape(p.shape), q, tau, t0, p.fm, p.sens, 0.0).reshape(-1)
           + orc.sparse_reg(v, lam, *coo))
    err = rel_l2(got, ref)
    print(f"3D toeplitz + sparse reg: {err:.2e}")
    assert err < 2e-6
    op.close()
