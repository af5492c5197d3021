"""toeplitz_recon's outer loops (toeplitz/recon.cu:2011-2133) through the
file-level entry point imp_toeplitz_recon: total-variation passes (W from
the current image before every pass after the first) and CG restarts (the
best residual carried from restart 1 on), against oracle.toeplitz_recon with
FP32 vector storage; and the context rules of the operators."""
import ctypes as C

import numpy as np
import pytest

from conftest import difference_coo, rel_l2

pytestmark = pytest.mark.gpu


def _recon_dir(ctx, tmp_path, Nz, ncoils, L, seed, scale=1.0):
    from impatient import files, ops, synth
    from oracle import oracle as orc
    p = synth.tiny(N=16, Nz=Nz, ncoils=ncoils, ntime_segs=L, seed=seed)
    p.rho = p.rho * scale
    d = orc.nudft_forward(p.kx, p.ky, p.kz, p.t, p.x, p.y, p.z, p.fm, p.rho, p.sens)
    d = synth.add_noise(d, 30.0, seed)
    files.write_problem(tmp_path, p, d)
    b = ops.nudft_adjoint(ctx, p.kx, p.ky, p.kz, p.t, p.x, p.y, p.z, p.fm, d, p.sens)
    q = np.asarray(ops.toeplitz_q(ctx, p.kx, p.ky, p.kz, p.t, p.Nx, p.Ny, Nz, L))
    return p, np.asarray(b).reshape(-1), q


def _run_recon(tmp_path, p, b, q, lam, iters, tv_num, enable_tv, restarts=1):
    from impatient import _lib
    lib = _lib.load()
    L = q.shape[0] - 1
    Qr = [np.ascontiguousarray(q[l].real, np.float32) for l in range(L + 1)]
    Qi = [np.ascontiguousarray(q[l].imag, np.float32) for l in range(L + 1)]
    PP = C.POINTER(C.c_float) * (L + 1)
    qr = PP(*[a.ctypes.data_as(C.POINTER(C.c_float)) for a in Qr])
    qi = PP(*[a.ctypes.data_as(C.POINTER(C.c_float)) for a in Qi])
    br = np.ascontiguousarray(b.real, np.float32)
    bi = np.ascontiguousarray(b.imag, np.float32)
    out = tmp_path / "out.file"
    argv = ["toeplitz_recon", str(p.Ny), str(p.Nx), str(p.Nz), str(restarts), str(iters), "0",
            f"{lam:f}", "D.file", "Dp.file", "R.file", "W.file", "Q_stone.file", "FhD.file",
            str(tmp_path), str(tv_num), f"{float(L):f}", str(out)]
    cargv = (C.c_char_p * 18)(*[a.encode() for a in argv])
    _lib.check(lib.imp_toeplitz_recon(18, cargv, int(enable_tv), 0, 1, 0, None,
                                      br.ctypes.data_as(C.c_void_p),
                                      bi.ctypes.data_as(C.c_void_p), qr, qi),
               "imp_toeplitz_recon")
    img = np.fromfile(out, dtype=np.float32).reshape(2, -1)
    return img[0] + 1j * img[1]


def _oracle_recon(p, b, q, lam, iters, tv_num, enable_tv, restarts=1, vec=np.complex64):
    from oracle import oracle as orc
    L = q.shape[0] - 1
    tau, t0 = orc.segment_timing(p.t, L)
    shape = (p.Ny, p.Nx, p.Nz) if p.Nz > 1 else (p.Ny, p.Nx)
    make = lambda w: orc.toeplitz_operator(q, tau, t0, p.shape, p.fm, p.sens, lam, w)
    x, hists = orc.toeplitz_recon(make, b.astype(np.complex128), shape, iters,
                                  enable_tv=enable_tv, tv_num=tv_num, restarts=restarts,
                                  vec_dtype=vec)
    return x, hists


# The TV weight w = 0.5 / sqrt(|D x|^2 + 1e-6) of the current image makes
# the later passes sensitive to rounding where |D x|^2 is near the 1e-6
# floor: on the unit-scale phantom below the oracle's own FP32-vector run
# leaves its FP64 run by 1e-3 (3D) to 8e-2 (2D) after three passes, while a
# single pass agrees to 1e-5.  Scaling the data by 1e3 (lambda 0.5) moves
# |D x|^2 well above the floor, and in 3D the oracle's FP32 and FP64 runs
# then agree to 3e-6: there the GPU is held to 1e-4.  In 2D every setting
# tried (N 16-24, 512-2048 samples, lambda 2e-3-5, SNR 30-60 dB) keeps a
# 1e-2-8e-2 FP32/FP64 spread: there the GPU is held to that spread.
TV_SCALED = (1e3, 0.5)


@pytest.mark.parametrize("Nz", [1, 4])
def test_toeplitz_recon_tv_passes_match_oracle(ctx, tmp_path, Nz):
    """-tv -tv_num 3 on the Toeplitz driver, 2D and 3D, unit-scale phantom:
    three passes of 5 iterations, each after the first with W = tv_weight of
    the current image, against the oracle's restatement of
    recon.cu:2011-2133 with FP32 vectors.  One pass is held to 1e-4; three
    passes to the oracle's own FP32-vs-FP64 spread (in 2D that spread,
    8.5e-2, is a third of what the TV passes change: there the later passes
    are not fixed by the arithmetic), and the GPU's distance is under a
    tenth of the TV passes' change of the image."""
    lam = 2e-3
    p, b, q = _recon_dir(ctx, tmp_path, Nz, 2, 3, seed=41 + Nz)
    x = _run_recon(tmp_path, p, b, q, lam, 5, 3, True)
    ref, hists = _oracle_recon(p, b, q, lam, 5, 3, True)
    ref64, _ = _oracle_recon(p, b, q, lam, 5, 3, True, vec=np.complex128)
    one = _run_recon(tmp_path, p, b, q, lam, 5, 3, False)
    ref1, _ = _oracle_recon(p, b, q, lam, 5, 1, False)
    e_tv, e_one, moved = rel_l2(x, ref), rel_l2(one, ref1), rel_l2(ref, ref1)
    spread = rel_l2(ref, ref64)
    print(f"Nz={Nz}: 3 TV passes vs FP32-storage oracle {e_tv:.2e} (oracle FP32 vs FP64 "
          f"{spread:.2e}); one pass {e_one:.2e}; TV moves the image by {moved:.2e}")
    assert len(hists) == 3
    assert e_one < 1e-4
    assert e_tv < max(1e-4, 2 * spread)
    assert moved > 10 * e_tv


def test_toeplitz_recon_tv_passes_match_oracle_well_conditioned(ctx, tmp_path):
    """The same three TV passes in 3D on data scaled by 1e3 (lambda 0.5),
    where the oracle's FP32 and FP64 runs agree to a few 1e-6: GPU against
    the FP32-storage oracle within 1e-4."""
    scale, lam = TV_SCALED
    p, b, q = _recon_dir(ctx, tmp_path, 4, 2, 3, seed=45, scale=scale)
    x = _run_recon(tmp_path, p, b, q, lam, 5, 3, True)
    ref, hists = _oracle_recon(p, b, q, lam, 5, 3, True)
    ref64, _ = _oracle_recon(p, b, q, lam, 5, 3, True, vec=np.complex128)
    ref1, _ = _oracle_recon(p, b, q, lam, 5, 1, False)
    e, spread, moved = rel_l2(x, ref), rel_l2(ref, ref64), rel_l2(ref, ref1)
    print(f"3D scaled: 3 TV passes vs FP32-storage oracle {e:.2e} (oracle FP32 vs FP64 "
          f"{spread:.2e}); TV moves the image by {moved:.2e}")
    assert len(hists) == 3
    assert e < 1e-4 and moved > 1e-2


def test_toeplitz_recon_restarts_match_oracle(ctx, tmp_path):
    """numRestarts = 2 with 2 TV passes (argv[4], reachable through the C ABI):
    restart 1 carries beta and the best residual (recon.cu:2043-2048); 3D
    scaled data as above, within 1e-4 of the FP32-storage oracle."""
    scale, lam = TV_SCALED
    p, b, q = _recon_dir(ctx, tmp_path, 4, 2, 3, seed=47, scale=scale)
    x = _run_recon(tmp_path, p, b, q, lam, 4, 2, True, restarts=2)
    ref, hists = _oracle_recon(p, b, q, lam, 4, 2, True, restarts=2)
    ref64, _ = _oracle_recon(p, b, q, lam, 4, 2, True, restarts=2, vec=np.complex128)
    one, _ = _oracle_recon(p, b, q, lam, 4, 2, True, restarts=1)
    e, spread = rel_l2(x, ref), rel_l2(ref, ref64)
    print(f"2 restarts x 2 TV passes vs FP32-storage oracle: {e:.2e} (oracle FP32 vs FP64 "
          f"{spread:.2e}; the second restart moves the image by {rel_l2(ref, one):.2e})")
    assert len(hists) == 4
    assert e < 1e-4


def test_sparse_reg_of_another_context_is_rejected(ctx):
    """A regularizer is used on its own context's stream and detached only by
    that context: an operator or brute-force solve on another context must
    refuse it (both contexts on device 0)."""
    from impatient import _lib, ops, synth
    from impatient._lib import Context, ptr
    from oracle import oracle as orc
    c2 = Context(0)
    try:
        rows, cols, I, J, V = difference_coo((16, 16))
        R2 = ops.SparseReg(c2, rows, cols, I, J, V)
        p = synth.tiny(N=16, ncoils=1, ntime_segs=2)
        q = np.asarray(ops.toeplitz_q(ctx, p.kx, p.ky, None, p.t, 16, 16, 1, 2))
        tau, t0 = orc.segment_timing(p.t, 2)
        op = ops.Toeplitz(ctx, p.shape, q, tau, t0, p.fm, None, 1e-3)
        with pytest.raises(_lib.ImpError) as e:
            op.set_sparse_reg(R2)
        assert e.value.status == -1
        R1 = ops.SparseReg(ctx, rows, cols, I, J, V)
        op.set_sparse_reg(R1)          # the operator's own context: accepted
        op.set_sparse_reg(None)
        a = [ctx.alloc((max(p.nk, 256),)) for _ in range(12)]
        with pytest.raises(_lib.ImpError) as e:
            _lib.call("imp_bruteforce_cg_ex", ctx.h, C.c_int(p.nk), C.c_int(256), C.c_int(1),
                      ptr(a[0]), ptr(a[1]), ptr(a[2]), ptr(a[3]), ptr(a[4]), None, ptr(a[5]),
                      ptr(a[6]), None, None, C.c_int(3), C.c_float(1e-3), R2.h, C.c_int(1),
                      ptr(a[7]), ptr(a[8]))
        assert e.value.status == -1
        R1.close()
        op.close()
        R2.close()
    finally:
        c2.close()


def test_operator_calls_after_context_destroy_fail_cleanly(ctx):
    """Destroying a context detaches its Toeplitz operators: apply and CG
    then return IMP_EINVAL instead of touching the destroyed stream."""
    from impatient import _lib, ops, synth
    from impatient._lib import Context, ptr
    from oracle import oracle as orc
    c2 = Context(0)
    p = synth.tiny(N=16, ncoils=1, ntime_segs=2)
    q = np.asarray(ops.toeplitz_q(c2, p.kx, p.ky, None, p.t, 16, 16, 1, 2))
    tau, t0 = orc.segment_timing(p.t, 2)
    op = ops.Toeplitz(c2, p.shape, q, tau, t0, p.fm, None, 1e-3)
    c2.close()
    a = [ctx.alloc((256,)) for _ in range(4)]
    with pytest.raises(_lib.ImpError) as e:
        _lib.call("imp_toep_apply", op.h, ptr(a[0]), ptr(a[1]), ptr(a[2]), ptr(a[3]))
    assert e.value.status == -1
    st = _lib.CGStats()
    with pytest.raises(_lib.ImpError) as e:
        _lib.call("imp_toep_cg", op.h, ptr(a[0]), ptr(a[1]), ptr(a[2]), ptr(a[3]), C.c_int(3),
                  C.c_double(0.0), C.c_int(0), C.byref(st))
    assert e.value.status == -1
    op.close()
