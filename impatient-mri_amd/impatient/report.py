"""Per-configuration measurement report: F^H d, Q_l, the Toeplitz apply
(per-pass HBM GB/s) and CG / Chan PCG to relative residual 1e-5 with FP64
scalars, on one GPU, for any of the synthetic workload configurations.

    python -m impatient.report --config 2

Prints progress to stderr and one JSON line to stdout.  Kernel times come
from the library's kernel timer (HIP events on the context stream); the
algorithmic bytes of the Toeplitz passes count each pass's reads and writes
once (complex64 = 8 bytes) on the pruned 2x grid:

    pass A  reads p, the field map and S (8 + 4 + 8C bytes per voxel), writes
            the G1 lines along the last axis (8C x 2 per voxel)
    pass B  (3D) reads the N2 kept x values, writes 2N2            (16C, 32C)
    pass C  reads and writes the N1 kept y values of every (x, z)
            line of the 2x grid, plus Qf_l (real for the conjugate-symmetric
            Q_l of the standard grid: 4 x 2^d per voxel)
    pass D  (3D) reads 2N2 x values, keeps N2                      (32C, 16C)
    pass E  reads the G1 lines and S, keeps N along the last axis

With --sparse-reg the report also times lambda C^T W C with C the periodic
difference matrix of the image (the -reg option's c.rmtx, 2 or 3 blocks of
N rows, 2 nonzeros per row) against the finite-difference stencil that
computes the same operator; algorithmic bytes per voxel and block: the two
CSR copies (4 + 4 x 2 + 4 x 2 bytes per row of C and of C^T), W, p and the
work vector Cf (8 bytes, written then read), and out (read and written).
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import ops, synth
from ._lib import (Context, IMP_CG_BEST_ITERATE, IMP_CG_PRECOND, IMP_CG_TRUE_RESIDUAL, call,
                   ptr)
import ctypes as C


def pass_bytes(nv, ncoils, d3):
    """Algorithmic bytes per launch of each Toeplitz pass (one segment, all
    coils)."""
    c = ncoils
    if not d3:
        half = 2 * nv                              # N1 x 2 N2 points
        return {"toep_pass_a": 8 * nv + 4 * nv + 8 * c * nv + 8 * c * half,
                "toep_pass_c": 8 * c * half * 2 + 4 * 4 * nv,
                "toep_pass_e": 8 * c * half + 8 * c * nv}
    return {"toep_pass_a": 8 * nv + 4 * nv + 8 * c * nv + 16 * c * nv,
            "toep_pass_b": 16 * c * nv + 32 * c * nv,
            "toep_pass_c": 2 * 32 * c * nv + 4 * 8 * nv,
            "toep_pass_d": 32 * c * nv + 16 * c * nv,
            "toep_pass_e": 16 * c * nv + 8 * c * nv}


def difference_matrix(shape):
    """[D_0; D_1; ...] with (D_b p)(i) = p(i) - p(i + e_b) periodic, COO."""
    import scipy.sparse as sp
    blocks = []
    for b, na in enumerate(shape):
        D = sp.identity(na, format="csr") - sp.csr_matrix(
            (np.ones(na), (np.arange(na), (np.arange(na) + 1) % na)), shape=(na, na))
        K = None
        for a, s in enumerate(shape):
            m = D if a == b else sp.identity(s, format="csr")
            K = m if K is None else sp.kron(K, m, format="csr")
        blocks.append(K)
    Cm = sp.vstack(blocks).tocoo()
    return Cm.shape[0], Cm.shape[1], Cm.row, Cm.col, Cm.data


def sparse_reg_timing(ctx, p, timer_on, kernel_ms, log):
    """lambda C^T W C p (C = difference matrix) vs the stencil, on the device."""
    shape = (p.Ny, p.Nx, p.Nz) if p.Nz > 1 else (p.Ny, p.Nx)
    rows, cols, I, J, V = difference_matrix(shape)
    R = ops.SparseReg(ctx, rows, cols, I, J, V)
    rng = np.random.default_rng(1)
    v = (rng.standard_normal(cols) + 1j * rng.standard_normal(cols)).astype(np.complex64)
    w = rng.uniform(0.5, 1.5, cols).astype(np.float32)
    a = R.apply(v, 1e-3, w)
    ref = ops.fd_reg(ctx, v.reshape(shape), 1e-3, w.reshape(shape)).reshape(-1)
    diff = float(np.linalg.norm(a - ref) / np.linalg.norm(ref))
    timer_on()
    for _ in range(5):
        R.apply(v, 1e-3, w)
        ops.fd_reg(ctx, v.reshape(shape), 1e-3, w.reshape(shape))
    res = {"rows": rows, "cols": cols, "nnz": int(len(V)), "rel_diff_vs_stencil": diff}
    nnz, nd = len(V), rows // cols
    # CSR of C (rows+1 pointers, nnz idx + val), Cf written, p gathered, W
    b_wc = 4 * (rows + 1) + 8 * nnz + 8 * rows + 8 * cols + 4 * cols
    # CSR of C^T, Cf gathered, out read and written
    b_t = 4 * (cols + 1) + 8 * nnz + 8 * rows + 16 * cols
    b_st = 8 * cols + 4 * cols + 16 * cols   # p, W, out read and written
    for name, nbytes in (("k_csr_wc", b_wc), ("k_csr_t", b_t), ("k_fd_reg", b_st)):
        tot, n = kernel_ms(name)
        avg = tot / max(n, 1)
        res[name] = {"avg_ms": round(avg, 4), "algorithmic_bytes": nbytes,
                     "gbs": round(nbytes / (avg * 1e-3) / 1e9, 1) if avg > 0 else None}
        log(f"{name}: {n} launches, {avg:.4f} ms, {res[name]['gbs']} GB/s")
    log(f"sparse reg ({nd} blocks, {nnz} nonzeros) vs stencil: rel diff {diff:.2e}")
    R.close()
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--cg-max-iter", type=int, default=1000)
    ap.add_argument("--tol", type=float, default=1e-5)
    ap.add_argument("--sparse-reg", action="store_true",
                    help="also time the sparse-matrix regularizer against the stencil")
    args = ap.parse_args(argv)

    def log(m):
        print(f"[report] {m}", file=sys.stderr, flush=True)

    ctx = Context(0)

    def timer_on():
        call("imp_profile_enable", ctx.h, C.c_int(1))
        call("imp_profile_reset", ctx.h)

    def kernel_ms(name):
        tot, n = C.c_double(), C.c_int()
        call("imp_profile_read", ctx.h, name.encode(), C.byref(tot), C.byref(n))
        return tot.value, n.value

    p = synth.config(args.config)
    d3 = p.Nz > 1
    nk, nv, nc, L = p.nk, p.nvox, p.ncoils, p.ntime_segs
    log(f"{p.name}: M={nk} N={nv} C={nc} L={L}")
    t0 = time.perf_counter()
    d = ops.nudft_forward(ctx, p.kx, p.ky, p.kz, p.t, p.x, p.y, p.z, p.fm, p.rho, p.sens, nc)
    d = synth.add_noise(np.asarray(d).reshape(nc, -1), 40.0, p.meta["seed"])
    log(f"data: forward NUDFT + noise {time.perf_counter() - t0:.2f} s")

    # F^H d: one warm-up, one timed by the kernel timer
    ops.nudft_adjoint(ctx, p.kx, p.ky, p.kz, p.t, p.x, p.y, p.z, p.fm, d, p.sens)
    timer_on()
    b = ops.nudft_adjoint(ctx, p.kx, p.ky, p.kz, p.t, p.x, p.y, p.z, p.fm, d, p.sens)
    # direct kernels, or the stacked path's group launches and its reduction
    parts = {k: kernel_ms(k) for k in ("k_nudft_mfma", "k_nudft_valu", "k_nudft_stack",
                                       "k_reduce_stack")}
    fhd_ms = sum(v[0] for v in parts.values())
    stacked = parts["k_nudft_stack"][1] > 0
    path = "stacked (z-separable)" if stacked else "direct"
    pairs = float(nc) * nk * nv
    # (sample, voxel) pairs whose phase and sin/cos were evaluated: once per
    # pair for all coils on the direct path; on the stacked path once per
    # pair of one partition and group launch of 32 coil-partition columns
    if stacked:
        K = nk // p.Nz
        evaluated = float(parts["k_nudft_stack"][1]) * K * nv
    else:
        evaluated = float(nk) * nv
    log(f"F^H d ({path}): {fhd_ms:.1f} ms, {pairs / fhd_ms / 1e6:.1f} GSample*vox/s; "
        + ", ".join(f"{k} {v[1]} x {v[0] / max(v[1], 1):.2f} ms" for k, v in parts.items() if v[1]))

    # Q_l and the plan from device arrays (no host round trip of the planes)
    up = lambda a: ctx.to_device(np.ascontiguousarray(a, dtype=np.float32).reshape(-1))
    dkx, dky, dt_ = up(p.kx), up(p.ky), up(p.t)
    dkz = up(p.kz) if d3 else None
    nq = (8 if d3 else 4) * nv
    qr, qi = ctx.alloc(((L + 1) * nq,)), ctx.alloc(((L + 1) * nq,))
    ctx.synchronize()
    t0 = time.perf_counter()
    call("imp_toeplitz_q", ctx.h, C.c_int(nk), ptr(dkx), ptr(dky), ptr(dkz), ptr(dt_), None,
         None, C.c_int(p.Nx), C.c_int(p.Ny), C.c_int(p.Nz), C.c_int(nq), None, None, None,
         C.c_int(L), C.c_int(0), ptr(qr), ptr(qi))
    ctx.synchronize()
    t_q = time.perf_counter() - t0
    log(f"Q_l: {t_q:.3f} s")
    t = p.t.astype(np.float32)
    tau = float(np.float32((t[-1] - t[0]) / np.float32(L)))
    t0 = time.perf_counter()
    op = ops.Toeplitz(ctx, p.shape, (qr, qi), tau, float(t[0]), p.fm, p.sens, p.lam)
    ctx.synchronize()
    t_plan = time.perf_counter() - t0
    op.apply(b)
    timer_on()
    for _ in range(3):
        op.apply(b)
    passes = {}
    for name, nbytes in pass_bytes(nv, nc, d3).items():
        tot, n = kernel_ms(name)
        avg = tot / max(n, 1)
        passes[name] = {"launches": n, "avg_ms": round(avg, 4), "algorithmic_bytes": nbytes,
                        "gbs": round(nbytes / (avg * 1e-3) / 1e9, 1) if avg > 0 else None}
        log(f"{name}: {n} launches, {avg:.4f} ms, {passes[name]['gbs']} GB/s")
    call("imp_profile_enable", ctx.h, C.c_int(0))
    sreg = None
    if args.sparse_reg:
        sreg = sparse_reg_timing(ctx, p, timer_on, kernel_ms, log)
    base = IMP_CG_BEST_ITERATE | IMP_CG_TRUE_RESIDUAL
    solves = {}
    for name, flags in (("plain", base), ("chan", base | IMP_CG_PRECOND)):
        _, st = op.cg(b, args.cg_max_iter, args.tol, flags)
        solves[name] = {"iterations": st["iterations"], "converged": st["converged"],
                        "seconds": round(st["seconds"], 4),
                        "true_rel_residual": st["true_rel_residual"],
                        "time_to_tol_s": round(fhd_ms * 1e-3 + t_q + t_plan + st["seconds"], 4)}
        log(f"{name} CG: {st['iterations']} it, {st['seconds']:.3f} s, "
            f"true residual {st['true_rel_residual']:.2e}")
    op.close()
    out = {"config": p.name, "nsamples": nk, "nvox": nv, "ncoils": nc, "segments": L,
           "fhd_path": path, "fhd_ms": round(fhd_ms, 3),
           "fhd_direct_equivalent_pairs": pairs, "fhd_evaluated_phase_pairs": evaluated,
           "fhd_gsample_vox_s": round(pairs / fhd_ms / 1e6, 1),
           "q_s": round(t_q, 4), "toeplitz_plan_s": round(t_plan, 4),
           "toeplitz_passes": passes, "tol": args.tol, "solves": solves}
    if sreg is not None:
        out["sparse_reg"] = sreg
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
