#!/usr/bin/env python3
"""Benchmark of the IMPATIENT hot path on one MI355X.

Workload: 512x512x1, 32 coils, variable-density spiral of 48 interleaves x
8192 samples (393216 samples), readout 0-10 ms, field map within +-200 Hz,
L = 10 time segments, lambda = 1e-3 (synthetic data: Shepp-Logan phantom,
Gaussian ring sensitivities, forward NUDFT + 40 dB noise).

A step is one direct, field-corrected F^H d over the whole workload
(32 coils x 393216 samples x 262144 voxels = 3.30e12 sample-voxel pairs),
inputs resident in HBM.  value = pairs / time in GSample*vox/s.
The same run also reports, once each and after one untimed call (the first
call of a kernel loads its code): the Toeplitz kernels Q_l, the resident
FFT(Q_l) plan, and plain CG, Chan PCG and the learned solve to relative
residual 1e-5.

The roofline entry is for the dominant kernel (k_nudft_mfma), from HIP
events on the library's stream around every step: 24 fp16 MFMA flops per
(coil, sample, voxel) pair (a complex MAC as three fp16 products) against the
dense fp16 MFMA peak.
The cpu_baseline leg times the C restatement of the reference CPU path
(iftCpu loop order, FP32 sinf/cosf, OpenMP inner reduction) on a bounded
sample of the same adjoint and converts it to the same unit.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "impatient-mri_amd"))

PEAK_FP16_MFMA_TFLOPS = 2500.0   # MI355X dense fp16 MFMA
PEAK_HBM_GBS = 8000.0
PEAK_FP32_VALU_TFLOPS = 157.3    # MI355X FP32 vector
VALU_FLOPS_PER_PAIR = 8.0        # one complex MAC in FP32
# fp16 MFMA flops per (coil, sample, voxel) pair: one complex MAC is 8 real
# flops, done as 3 fp16 products (Eh Wh + Eh Wl + El Wh) for FP32 accuracy
MFMA_FLOPS_PER_PAIR = 24.0
# HBM bytes per F^H d launch of k_nudft_mfma<2,false> on this workload, from
# rocprofv3 --pmc FETCH_SIZE (x2.0) and WRITE_SIZE (x1.0), the factors a
# 256 MiB device copy calibrates (profiles/traffic_512x512x1_32coil_pmc.md);
# re-measure when the kernel changes
NUDFT_TRAFFIC_BYTES = 2.0 * 3414246.9 * 1024 + 1.0 * 957937.5 * 1024
NUDFT_TRAFFIC_SOURCE = ("constant from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of "
                        "bench.py --pmc-mode (profiles/traffic_512x512x1_32coil_pmc.md), "
                        "not measured in this run")


def hip_events():
    """hipEvent helpers on a given stream (libamdhip64 via ctypes)."""
    lib = C.CDLL("libamdhip64.so")

    class Ev:
        def __init__(self):
            self.h = C.c_void_p()
            assert lib.hipEventCreate(C.byref(self.h)) == 0

        def record(self, stream):
            assert lib.hipEventRecord(self.h, C.c_void_p(stream)) == 0

        def sync(self):
            assert lib.hipEventSynchronize(self.h) == 0

        def ms_since(self, start):
            v = C.c_float()
            assert lib.hipEventElapsedTime(C.byref(v), start.h, self.h) == 0
            return v.value
    return Ev


def cpu_baseline(prob, d_host, seconds_target=15.0):
    """Time the oracle's C restatement of iftCpu + parallelIftCpu on a bounded
    sample: 1 coil, the first `m` samples, every voxel."""
    from oracle import oracle as _orc  # noqa: F401  (oracle package: checker/baseline only)
    so = REPO / "oracle" / "liboracle_cpu.so"
    if not so.exists():
        return None
    lib = C.CDLL(str(so))
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    nv = prob.nvox
    out_r = np.zeros(nv, np.float32)
    out_i = np.zeros(nv, np.float32)
    kx, ky, t = f32(prob.kx), f32(prob.ky), f32(prob.t)
    x, y, fm = f32(prob.x), f32(prob.y), f32(prob.fm)
    dr, di = f32(d_host[0].real), f32(d_host[0].imag)
    # The restated loop has a per-voxel parallel region over samples, so its
    # time is a + b*m rather than b*m: fit both from two samples, then size
    # the timed sample to about seconds_target.
    def run(m):
        t0 = time.perf_counter()
        lib.ift_cpu(P(out_r), P(out_i), P(dr), P(di), P(kx), P(ky), P(x), P(y), P(fm), P(t),
                    C.c_int(m), C.c_int(nv))
        return time.perf_counter() - t0
    run(16)  # thread pool start-up
    m1, m2 = 128, 512
    t1, t2 = run(m1), run(m2)
    b = max((t2 - t1) / (m2 - m1), 1e-9)
    a = max(t1 - b * m1, 0.0)
    m = int(min(prob.nk, max(m2, (seconds_target - a) / b)))
    dt = run(m)
    if dt < 0.6 * seconds_target and m < prob.nk:   # small samples overstate b: refine once
        m = int(min(prob.nk, m * seconds_target / dt))
        dt = run(m)
    pairs = float(m) * nv
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": pairs / dt / 1e9, "unit": "GSample*vox/s", "cores": cores,
            "kind": "port", **host_cpu(),
            "sample": f"iftCpu restatement (FP32, OpenMP inner reduction), 1 coil x {m} "
                      f"samples x {nv} voxels = {pairs:.3e} pairs in {dt:.2f} s"}


def host_cpu():
    """CPU model (from /proc/cpuinfo) and the logical CPUs this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": avail}


def cpu_vector_ops(prob, fd_gpu_ms):
    """The CG's vector stages at the image size, on the host (the C
    restatements of DHWD2dCpu, dotProductCpu and addCpu, one thread as in the
    reference) beside the GPU's kernels from the bench's own kernel timer:
    seconds per call and algorithmic GB/s (stencil: p read, s written; dot:
    two vectors read; axpy: x read, y read and written; complex64 = 8 B)."""
    from oracle import oracle as _orc  # noqa: F401  (oracle package: checker/baseline only)
    so = REPO / "oracle" / "liboracle_cpu.so"
    if not so.exists():
        return None
    lib = C.CDLL(str(so))
    lib.dot_re_cpu.restype = C.c_float
    n = prob.nvox
    rng = np.random.default_rng(3)
    a = [np.ascontiguousarray(rng.standard_normal(n), np.float32) for _ in range(4)]
    o = [np.zeros(n, np.float32) for _ in range(2)]
    P = lambda v: v.ctypes.data_as(C.POINTER(C.c_float))
    calls = {"fd_reg": lambda: lib.fd_reg_cpu(P(o[0]), P(o[1]), P(a[0]), P(a[1]), prob.Nx,
                                              prob.Ny, C.c_float(1e-3)),
             "dot": lambda: lib.dot_re_cpu(P(a[0]), P(a[1]), P(a[2]), P(a[3]), n),
             "axpy": lambda: lib.axpy_cpu(P(o[0]), P(o[1]), C.c_float(0.5), P(a[0]), P(a[1]), n)}
    nbytes = {"fd_reg": 16 * n, "dot": 16 * n, "axpy": 24 * n}
    out = {}
    for name, fn in calls.items():
        fn()
        reps = 20
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        dt = (time.perf_counter() - t0) / reps
        out[name] = {"cpu_s": round(dt, 6), "cpu_gbs": round(nbytes[name] / dt / 1e9, 2)}
    if fd_gpu_ms:
        out["fd_reg"]["gpu_kernel"] = "k_fd_reg"
        out["fd_reg"]["gpu_s"] = round(fd_gpu_ms * 1e-3, 7)
    out["note"] = "one image of the 512^2 workload per call"
    out["threads"] = 1
    return out


def cpu_bruteforce_cfg0(ctx):
    """The reference's CPU brute force on its own small case (64x64, 1 coil,
    8 x 1024 spiral, cg_num = 10, lambda = 1e-3): the oracle's FP32 C
    restatement of bruteForceCpu (bruteforce_cpu) timed in full on the host,
    beside imp_bruteforce_cg on the GPU (wall time from host buffers)."""
    from oracle import oracle as _orc  # noqa: F401  (oracle package: checker/baseline only)
    from impatient import ops, synth
    so = REPO / "oracle" / "liboracle_cpu.so"
    if not so.exists():
        return None
    lib = C.CDLL(str(so))
    p = synth.config(0)
    d = ops.nudft_forward(ctx, p.kx, p.ky, None, p.t, p.x, p.y, None, p.fm, p.rho, None, 1)
    d = synth.add_noise(np.asarray(d).reshape(1, -1), 40.0, p.meta["seed"])
    x0 = np.zeros(p.nvox, np.complex64)
    run_gpu = lambda: ops.bruteforce_cg(ctx, p.kx, p.ky, p.t, p.x, p.y, p.fm, d, None,
                                        p.cg_iters, p.lam, x0)
    run_gpu()                                   # untimed: kernel code loading
    t0 = time.perf_counter()
    xg = run_gpu()
    t_gpu = time.perf_counter() - t0
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    keep = [f32(v) for v in (p.kx, p.ky, p.x, p.y, p.fm, p.t, d.real.reshape(-1),
                             d.imag.reshape(-1))]
    xr = np.zeros(p.nvox, np.float32)
    xi = np.zeros(p.nvox, np.float32)
    t0 = time.perf_counter()
    lib.bruteforce_cpu(P(xr), P(xi), P(keep[6]), P(keep[7]), P(keep[0]), P(keep[1]), P(keep[2]),
                       P(keep[3]), P(keep[4]), P(keep[5]), None, None, C.c_int(1), C.c_int(p.nk),
                       C.c_int(p.Nx), C.c_int(p.Ny), C.c_int(p.cg_iters), C.c_float(p.lam))
    t_cpu = time.perf_counter() - t0
    xc = xr + 1j * xi
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"workload": "64x64x1 1-coil spiral 8x1024, brute-force CG, 9 iterations "
                        "(cg_num 10), lambda 1e-3",
            "cpu_s": round(t_cpu, 4), "gpu_s": round(t_gpu, 4),
            "gpu_speedup": round(t_cpu / t_gpu, 1), "cores": cores, "kind": "port",
            "gpu_vs_cpu_rel_l2": float(np.linalg.norm(xg - xc) / np.linalg.norm(xc))}


MODEL = REPO / "impatient-mri_amd" / "models" / "cfg3_warmstart.npz"


def learned_solve(args, ctx, h, prob, dbr, dbi, dxr, dxi, fhd_ms, t_q, t_plan, log):
    """Learned warm start + learned-gain circulant PCG to 1e-5 (the trained
    model of impatient.learn), timed from b = F^H d on the device: z0 =
    M_chan^-1 b, x0 = z0 + s * net(z0 / s), PCG from x0 (zero start if x0's
    residual is worse)."""
    import impatient
    from impatient import ops, learn
    from impatient._lib import call, ptr
    if not MODEL.exists():
        log(f"no trained model at {MODEL}: learned start skipped")
        return f"no model file ({MODEL.name})"
    params, gains, meta = learn.load_model(MODEL)
    params_pc = learn.load_precond_params(MODEL) if meta.get("use_netw") else None
    net = ops.CNN(ctx, prob.Ny, prob.Nx, seed=1)
    net.set_params(params)
    gh = np.ascontiguousarray(gains, np.float32)
    # the preconditioner network's weights depend on the trajectory (through
    # the Chan weights of Q_l), not on the data: computed once per
    # trajectory, like Q_l, and counted in the time to 1e-5
    dw = None
    t_pcnet = 0.0
    if params_pc is not None:
        net_pc = ops.CNN(ctx, prob.Ny, prob.Nx, seed=1)
        net_pc.set_params(params_pc)
        for _ in range(2):   # the first call loads the kernels' code; the second is timed
            call("imp_toep_set_precond", h, None)
            ctx.synchronize()
            t0 = time.perf_counter()
            chan = ctx.alloc((prob.nvox,))
            call("imp_toep_get_precond", h, ptr(chan))
            w = learn.network_weights(net_pc, chan.get().reshape(prob.Ny, prob.Nx))
            dw = ctx.to_device(np.ascontiguousarray(w.reshape(-1), np.float32))
            ctx.synchronize()
            t_pcnet = time.perf_counter() - t0
        net_pc.close()
    flags = (impatient.IMP_CG_BEST_ITERATE | impatient.IMP_CG_TRUE_RESIDUAL |
             impatient.IMP_CG_PRECOND | impatient.IMP_CG_USE_X0 | impatient.IMP_CG_FALLBACK_ZERO)
    zr, zi = ctx.alloc((prob.nvox,)), ctx.alloc((prob.nvox,))
    out = {"model": MODEL.name, "trained_on": f"{meta['draws']} draws, {meta['epochs']} epochs",
           "preconditioner": ("network weights" if dw is not None else
                              "learned radial gains" if meta.get("use_gains") else "Chan")}
    for warm in (True, False):   # first pass warms the kernels up, second is timed
        call("imp_toep_set_precond", h, None)
        ctx.synchronize()
        t0 = time.perf_counter()
        call("imp_toep_precond_apply", h, ptr(dbr), ptr(dbi), ptr(zr), ptr(zi))
        net.forward_device(zr, zi, dxr, dxi)
        if dw is not None:
            call("imp_toep_set_precond", h, ptr(dw))
        elif meta.get("use_gains", False):
            call("imp_toep_set_precond_radial", h, gh.ctypes.data_as(C.c_void_p),
                 C.c_int(len(gh)))
        ctx.synchronize()
        t_start = time.perf_counter() - t0
        st = _lib_stats()
        call("imp_toep_cg", h, ptr(dbr), ptr(dbi), ptr(dxr), ptr(dxi), C.c_int(args.cg_max_iter),
             C.c_double(1e-5), C.c_int(flags), C.byref(st))
        t_total = time.perf_counter() - t0
    s = st.as_dict()
    log(f"learned start + PCG ({out['preconditioner']}): start {t_start * 1e3:.1f} ms (residual "
        f"{s['start_rel_residual']:.3e}, used {s['start_used']}), {s['iterations']} it, "
        f"{t_total:.3f} s total, true {s['true_rel_residual']:.2e}")
    call("imp_toep_set_precond", h, None)
    net.close()
    out.update({"iterations": s["iterations"], "start_s": round(t_start, 4),
                "start_rel_residual": s["start_rel_residual"], "start_used": s["start_used"],
                "solve_s": round(t_total, 4), "true_rel_residual": s["true_rel_residual"],
                "converged": s["converged"],
                "precond_network_s": round(t_pcnet, 4),
                "time_to_1e-5_s": round(fhd_ms * 1e-3 + t_q + t_plan + t_pcnet + t_total, 4)})
    return out


def batch_leg(ctx, log):
    """The 48-slice batch of uneven sizes ({128, 256, 384}^2, 8 coils, L = 8,
    +-150 Hz): one persistent F^H d launch, Q_l per size class, per slice the
    learned start (plain-CG fallback) and graph-captured PCG to 1e-5.  Two
    passes with one autotuner; the second is reported (the first includes the
    autotuner's trial launches)."""
    from impatient import batch, learn
    t0 = time.perf_counter()
    sl = batch.make_slices(48, ctx=ctx)
    t_syn = time.perf_counter() - t0
    params = learn.load_model(MODEL)[0] if MODEL.exists() else None
    tuner = batch.Autotuner(ctx)
    quiet = lambda m: None
    first = batch.reconstruct(ctx, sl, tol=1e-5, model=params, tuner=tuner, log=quiet)
    res = batch.reconstruct(ctx, sl, tol=1e-5, model=params, tuner=tuner, log=quiet)
    recs = res["slices"]
    by = {}
    for r in recs:
        by.setdefault(r["N"], []).append(r)
    out = {
        "workload": "48 2D slices from {128^2, 256^2, 384^2} (seeded draw), 8 coils, L=8, "
                    "+-150 Hz, spiral scaled per size; learned start with plain-CG fallback, "
                    "graph-captured PCG to 1e-5",
        "slices": len(recs), "synthesis_s": round(t_syn, 2),
        "first_pass_s": round(first["total_s"], 4), "total_s": round(res["total_s"], 4),
        "slices_per_s": round(len(recs) / res["total_s"], 2),
        "fhd_batch_s": round(res["fhd_s"], 4), "q_s": round(res["q_s"], 4),
        "solve_s": round(res["solve_s"], 4), "streams": res["streams"],
        "converged": int(sum(r["converged"] for r in recs)),
        "max_true_rel_residual": max(r["true_rel_residual"] for r in recs),
        "model": MODEL.name if params is not None else None,
        "per_size": {str(N): {"count": len(v),
                              "mean_iterations": round(float(np.mean([r["iterations"] for r in v])), 1),
                              "mean_slice_s": round(float(np.mean([r["slice_s"] for r in v])), 4),
                              "start_used": int(sum(r["start_used"] for r in v))}
                     for N, v in sorted(by.items())},
    }
    log(f"batch: {out['slices']} slices in {out['total_s']:.3f} s ({out['slices_per_s']} "
        f"slices/s), max true residual {out['max_true_rel_residual']:.2e}")
    return out


def gridding_leg(ctx, prob, dkx, dky, dt_, dfm, ddr, ddi, dsr, dsi, d_host, log):
    """The gridding path (-toeplitzGridding) on the same workload at the
    reference's demo oversampling (gridOS_Q 1.25, gridOS_FH 1.375;
    run_all_demos.sh:4): Q_l (L + 1 planes on the 1024^2 Q grid) and the
    32-coil F^H d on the GPU, each timed after one untimed call, beside the
    C restatement of the reference's CPU gold gridding step
    (gridding_Gold_2D, CPU_kernels.cpp:89-200, single-threaded as there):
    Q's 11 segment grids in full, F^H d's grids for 1 coil (x 32 coils,
    extrapolated)."""
    from impatient._lib import call, ptr
    from oracle import oracle as orc  # checker / baseline only
    L, nv, nk, C_ = prob.ntime_segs, prob.nvox, prob.nk, prob.ncoils
    nq = 4 * nv
    qr, qi = ctx.alloc(((L + 1) * nq,)), ctx.alloc(((L + 1) * nq,))
    br, bi = ctx.alloc((nv,)), ctx.alloc((nv,))

    def q():
        call("imp_toeplitz_q_gridding", ctx.h, C.c_int(nk), ptr(dkx), ptr(dky), None, ptr(dt_),
             C.c_int(prob.Nx), C.c_int(prob.Ny), C.c_int(1), C.c_int(L), C.c_float(1.25),
             ptr(qr), ptr(qi))

    def fhd():
        call("imp_fhd_gridding", ctx.h, C.c_int(nk), C.c_int(C_), ptr(dkx), ptr(dky), None,
             ptr(dt_), C.c_int(prob.Nx), C.c_int(prob.Ny), C.c_int(1), ptr(dfm), ptr(ddr),
             ptr(ddi), ptr(dsr), ptr(dsi), C.c_int(L), C.c_float(1.375), ptr(br), ptr(bi))
    out = {"gridOS_Q": 1.25, "gridOS_FH": 1.375}
    for name, fn in (("q", q), ("fhd", fhd)):
        fn()
        ctx.synchronize()
        t0 = time.perf_counter()
        fn()
        ctx.synchronize()
        out[name + "_s"] = round(time.perf_counter() - t0, 4)
    log(f"gridding on the GPU: Q_l {out['q_s'] * 1e3:.1f} ms, F^H d {out['fhd_s'] * 1e3:.1f} ms")
    so = REPO / "oracle" / "liboracle_cpu.so"
    if so.exists():
        lib = C.CDLL(str(so))
        P = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        f32 = lambda a: np.ascontiguousarray(a, np.float32)
        kx, ky, t = f32(prob.kx), f32(prob.ky), f32(prob.t)
        tau, t0 = orc.segment_timing(prob.t, L)
        lib.gridding_gold.argtypes = [C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                      C.c_void_p, C.c_float, C.c_void_p, C.c_void_p,
                                      C.POINTER(C.c_float), C.c_float, C.c_float, C.c_int,
                                      C.c_int, C.c_int, C.c_float, C.c_float,
                                      C.POINTER(C.c_int), C.POINTER(C.c_float),
                                      C.POINTER(C.c_float)]
        timings = {}
        for name, img, kscale, os_, data in (("q", (2 * prob.Nx, 2 * prob.Ny, 1), 2.0, 1.25, None),
                                            ("fhd_1coil", (prob.Nx, prob.Ny, 1), 1.0, 1.375,
                                             (f32(d_host[0].real), f32(d_host[0].imag)))):
            o, beta, g = orc.gridding_params(img, os_)
            gi = np.asarray(g, np.int32)
            gr_ = np.zeros(int(np.prod(g)), np.float32)
            gim = np.zeros_like(gr_)
            t1 = time.perf_counter()
            for l in range(L + 1):
                lib.gridding_gold(nk, P(kx), P(ky), None, kscale,
                                  None if data is None else data[0].ctypes.data,
                                  None if data is None else data[1].ctypes.data, P(t), float(l),
                                  float(tau), img[0], img[1], 1, float(o), float(beta),
                                  gi.ctypes.data_as(C.POINTER(C.c_int)), P(gr_), P(gim))
            timings[name] = time.perf_counter() - t1
        out["cpu_gold"] = {
            "what": "gridding step of gridding_Gold_2D (C restatement, 1 thread), all L+1 "
                    "segments; no FFT or deapodization (the reference runs those on the GPU)",
            "q_s": round(timings["q"], 3),
            "fhd_1coil_s": round(timings["fhd_1coil"], 3),
            "fhd_32coil_s_extrapolated": round(timings["fhd_1coil"] * C_, 2),
            "threads": 1}
        log(f"gridding, CPU gold step: Q {timings['q']:.2f} s, F^H d {timings['fhd_1coil']:.2f} s "
            f"per coil")
    return out


def _lib_stats():
    from impatient import _lib
    return _lib.CGStats()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-recon", action="store_true", help="skip Q / CG (profiling runs)")
    ap.add_argument("--cg-max-iter", type=int, default=400)
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 48-slice uneven-size batch")
    ap.add_argument("--pmc-mode", action="store_true",
                    help="counter runs: one F^H d, one Toeplitz apply and one 256 MiB "
                         "device copy (the calibration kernel), nothing else")
    args = ap.parse_args()
    if args.pmc_mode:
        args.steps, args.warmup, args.no_cpu_baseline = 1, 0, True
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path shards over coils with one image all-reduce per CG iteration; "
              "see the design notes).", file=sys.stderr)
        sys.exit(2)

    import impatient
    from impatient import ops, synth, _lib
    from impatient._lib import ptr, call

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    ctx = impatient.Context(0)
    stream = ctx.stream()
    Ev = hip_events()
    prob = synth.config(3)
    L = prob.ntime_segs
    nk, nv, ncoils = prob.nk, prob.nvox, prob.ncoils
    log(f"workload {prob.name}: M={nk} N={nv} C={ncoils} L={L}")

    up = lambda a: ctx.to_device(np.ascontiguousarray(a, dtype=np.float32))
    dkx, dky, dt_, dx, dy, dfm = (up(a) for a in (prob.kx, prob.ky, prob.t, prob.x, prob.y,
                                                  prob.fm))
    dsr, dsi = up(prob.sens.real.reshape(-1)), up(prob.sens.imag.reshape(-1))
    drr, dri = up(prob.rho.real), up(prob.rho.imag)
    ddr, ddi = ctx.alloc((ncoils * nk,)), ctx.alloc((ncoils * nk,))
    # synthetic k-space data: forward NUDFT of the phantom on the GPU + noise
    t0 = time.perf_counter()
    call("imp_nudft_forward", ctx.h, C.c_int(nk), C.c_int(nv), C.c_int(ncoils), ptr(dkx),
         ptr(dky), None, ptr(dt_), ptr(dx), ptr(dy), None, ptr(dfm), ptr(drr), ptr(dri),
         ptr(dsr), ptr(dsi), ptr(ddr), ptr(ddi))
    ctx.synchronize()
    t_fwd = time.perf_counter() - t0
    d = (ddr.get() + 1j * ddi.get()).reshape(ncoils, nk)
    d = synth.add_noise(d, 40.0, prob.meta["seed"])
    ddr.upload(d.real.reshape(-1))
    ddi.upload(d.imag.reshape(-1))
    log(f"data: forward NUDFT {t_fwd:.3f} s")

    dbr, dbi = ctx.alloc((nv,)), ctx.alloc((nv,))

    def fhd():
        call("imp_nudft_adjoint", ctx.h, C.c_int(nk), C.c_int(nv), C.c_int(ncoils), ptr(dkx),
             ptr(dky), None, ptr(dt_), ptr(dx), ptr(dy), None, ptr(dfm), None, None, ptr(ddr),
             ptr(ddi), ptr(dsr), ptr(dsi), ptr(dbr), ptr(dbi))

    def kernel_ms(name):
        tot, n = C.c_double(), C.c_int()
        call("imp_profile_read", ctx.h, name.encode(), C.byref(tot), C.byref(n))
        return tot.value, n.value

    for _ in range(args.warmup):
        fhd()
    ctx.synchronize()
    # library kernel timer: HIP events on the library stream around each
    # NUDFT / Toeplitz-pass launch inside the timed region
    call("imp_profile_enable", ctx.h, C.c_int(1))
    call("imp_profile_reset", ctx.h)
    evs = [Ev() for _ in range(args.steps + 1)]
    evs[0].record(stream)
    t0 = time.perf_counter()
    for i in range(args.steps):
        fhd()
        evs[i + 1].record(stream)
    evs[-1].sync()
    wall = time.perf_counter() - t0
    step_ms = [evs[i + 1].ms_since(evs[i]) for i in range(args.steps)]
    ms = float(np.mean(step_ms))
    pairs = float(ncoils) * nk * nv
    value = pairs / (ms * 1e-3) / 1e9
    log(f"F^H d: {ms:.2f} ms/step (events), wall {wall / args.steps * 1e3:.2f} ms/step, "
        f"{value:.1f} GSample*vox/s")

    k_tot, k_n = kernel_ms("k_nudft_mfma")
    k_ms = k_tot / max(k_n, 1)
    log(f"k_nudft_mfma: {k_n} launches, {k_ms:.2f} ms average")
    flops = MFMA_FLOPS_PER_PAIR * pairs
    achieved = flops / (k_ms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": round(achieved, 1), "peak": PEAK_FP16_MFMA_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP16_MFMA_TFLOPS, 4),
                "traffic": round(NUDFT_TRAFFIC_BYTES),
                "traffic_source": NUDFT_TRAFFIC_SOURCE,
                "kernel": "k_nudft_mfma<2,false> (F^H d, 32 coils in one pass)",
                "kernel_ms": round(k_ms, 3),
                "work": "24 fp16 MFMA flops per (coil, sample, voxel) pair"}
    # the same launch priced as an FP32 VALU NUDFT: one complex MAC (8 FP32
    # flops) per (coil, sample, voxel) pair, against the 157.3 TFLOP/s FP32
    # vector peak (the phase and sincos, shared by the coils, not counted)
    valu_eq = VALU_FLOPS_PER_PAIR * pairs / (k_ms * 1e-3) / 1e12
    roofline["fp32_valu_equivalent"] = {
        "flops_per_pair": VALU_FLOPS_PER_PAIR, "tflops": round(valu_eq, 1),
        "peak": PEAK_FP32_VALU_TFLOPS, "frac": round(valu_eq / PEAK_FP32_VALU_TFLOPS, 3)}

    result = {
        "metric": "NUDFT GSample*vox/s (direct field-corrected F^H d); time to rel-residual 1e-5 "
                  "(plain CG) reported in recon",
        "value": round(value, 2), "unit": "GSample*vox/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32-accurate (3x fp16-split MFMA products, FP32 accumulate, FP64 phase)",
        "data": "synthetic (seeded Shepp-Logan, spiral, Gaussian field map and coils)",
        "config": {"workload": "512x512x1 32-coil spiral 48x8192, L=10, +-200 Hz (F^H d)",
                   "nvox": nv, "nsamples": nk, "ncoils": ncoils, "pairs": pairs},
        "roofline": roofline,
    }

    if not args.no_recon:
        # Toeplitz kernels Q_l (L+1 planes on the 1024^2 grid)
        nq = 4 * nv
        dqr, dqi = ctx.alloc(((L + 1) * nq,)), ctx.alloc(((L + 1) * nq,))

        def q_call():
            call("imp_toeplitz_q", ctx.h, C.c_int(nk), ptr(dkx), ptr(dky), None, ptr(dt_), None,
                 None, C.c_int(prob.Nx), C.c_int(prob.Ny), C.c_int(1), C.c_int(nq), None, None,
                 None, C.c_int(L), C.c_int(0), ptr(dqr), ptr(dqi))
        # one untimed call first (kernel code loading, workspaces), as for F^H d
        q_call()
        ctx.synchronize()
        e0, e1 = Ev(), Ev()
        e0.record(stream)
        q_call()
        e1.record(stream)
        e1.sync()
        t_q = e1.ms_since(e0) * 1e-3
        q_pairs = float(L + 1) * nk * nq
        log(f"Q: {t_q:.3f} s ({q_pairs / t_q / 1e9:.1f} GSample*vox/s direct-equivalent)")
        t = prob.t.astype(np.float32)
        tau = float(np.float32((t[-1] - t[0]) / np.float32(L)))
        t0s = float(t[0])
        h = C.c_void_p()

        def plan():
            call("imp_toep_create", ctx.h, C.c_int(prob.Ny), C.c_int(prob.Nx), C.c_int(1),
                 C.c_int(ncoils), C.c_int(L), C.c_float(tau), C.c_float(t0s), ptr(dqr),
                 ptr(dqi), ptr(dfm), ptr(dsr), ptr(dsi), C.c_float(prob.lam), C.byref(h))
        plan()                          # untimed: the first plan loads its kernels
        call("imp_toep_destroy", h)
        tp0 = time.perf_counter()
        plan()
        t_plan = time.perf_counter() - tp0
        # the apply, timed with the kernel timer off (its per-launch events
        # would be inside the region), then 3 more with it on for the passes
        dxr, dxi = ctx.alloc((nv,)), ctx.alloc((nv,))
        call("imp_profile_enable", ctx.h, C.c_int(0))
        call("imp_toep_apply", h, ptr(dbr), ptr(dbi), ptr(dxr), ptr(dxi))
        ctx.synchronize()
        e0.record(stream)
        for _ in range(10):
            call("imp_toep_apply", h, ptr(dbr), ptr(dbi), ptr(dxr), ptr(dxi))
        e1.record(stream)
        e1.sync()
        t_apply = e1.ms_since(e0) / 10
        call("imp_profile_enable", ctx.h, C.c_int(1))
        call("imp_profile_reset", ctx.h)
        for _ in range(3):
            call("imp_toep_apply", h, ptr(dbr), ptr(dbi), ptr(dxr), ptr(dxi))
        ctx.synchronize()
        log(f"Toeplitz plan {t_plan:.3f} s, apply {t_apply:.2f} ms (timer off)")
        # HBM roofline of the FFT passes: algorithmic bytes per launch (one
        # segment l, all coils; complex64 = 8 B; the padded grid keeps the
        # N1 x 2N2 rows that hold data; Qf_l of this standard-grid Q_l is
        # real, 4 B per point of the 2x grid)
        half = 2 * nv                                   # N1 x 2 N2 points
        pass_bytes = {
            "toep_pass_a": 8 * nv + 4 * nv + 8 * ncoils * nv + 8 * ncoils * half,
            "toep_pass_c": 8 * ncoils * half * 2 + 4 * 4 * nv,
            "toep_pass_e": 8 * ncoils * half + 8 * ncoils * nv,
        }
        passes = {}
        for name, nbytes in pass_bytes.items():
            tot, n = kernel_ms(name)
            avg = tot / max(n, 1)
            passes[name] = {"launches": n, "avg_ms": round(avg, 4),
                            "algorithmic_bytes": nbytes,
                            "gbs": round(nbytes / (avg * 1e-3) / 1e9, 1) if avg > 0 else None}
            log(f"{name}: {n} launches, {avg:.4f} ms avg, {passes[name]['gbs']} GB/s")
        fd_tot, fd_n = kernel_ms("k_fd_reg")
        fd_gpu_ms = fd_tot / fd_n if fd_n else None
        call("imp_profile_enable", ctx.h, C.c_int(0))
        if args.pmc_mode:
            # calibration kernel for the byte counters: a 256 MiB device copy
            hip = C.CDLL("libamdhip64.so")
            nbytes = 256 << 20
            src, dst = ctx.alloc((nbytes // 4,)), ctx.alloc((nbytes // 4,))
            ctx.synchronize()
            assert hip.hipMemcpyAsync(C.c_void_p(dst.ptr), C.c_void_p(src.ptr),
                                      C.c_size_t(nbytes), C.c_int(3), C.c_void_p(stream)) == 0
            ctx.synchronize()
            call("imp_toep_destroy", h)
            log("pmc mode: F^H d, Q, 14 applies, one 256 MiB copy done")
            print(json.dumps(result), flush=True)
            return
        st = _lib.CGStats()
        tc0 = time.perf_counter()
        call("imp_toep_cg", h, ptr(dbr), ptr(dbi), ptr(dxr), ptr(dxi), C.c_int(args.cg_max_iter),
             C.c_double(1e-5), C.c_int(impatient.IMP_CG_BEST_ITERATE |
                                       impatient.IMP_CG_TRUE_RESIDUAL), C.byref(st))
        t_cg = time.perf_counter() - tc0
        s = st.as_dict()
        log(f"plain CG: {s['iterations']} it, {t_cg:.3f} s, rel res {s['rel_residual']:.2e}, "
            f"true {s['true_rel_residual']:.2e}")
        # preconditioned CG with the circulant (T. Chan) preconditioner
        st2 = _lib.CGStats()
        tc0 = time.perf_counter()
        call("imp_toep_cg", h, ptr(dbr), ptr(dbi), ptr(dxr), ptr(dxi), C.c_int(args.cg_max_iter),
             C.c_double(1e-5), C.c_int(impatient.IMP_CG_BEST_ITERATE |
                                       impatient.IMP_CG_TRUE_RESIDUAL | impatient.IMP_CG_PRECOND),
             C.byref(st2))
        t_pcg = time.perf_counter() - tc0
        s2 = st2.as_dict()
        log(f"Chan-preconditioned CG: {s2['iterations']} it, {t_pcg:.3f} s, rel res "
            f"{s2['rel_residual']:.2e}, true {s2['true_rel_residual']:.2e}")
        learned = learned_solve(args, ctx, h, prob, dbr, dbi, dxr, dxi, ms, t_q, t_plan, log)
        call("imp_toep_destroy", h)
        gridding = gridding_leg(ctx, prob, dkx, dky, dt_, dfm, ddr, ddi, dsr, dsi, d, log)
        result["recon"] = {
            "fhd_s": round(ms * 1e-3, 4), "q_s": round(t_q, 4),
            "q_direct_equiv_gsample_vox_s": round(q_pairs / t_q / 1e9, 1),
            "toeplitz_plan_s": round(t_plan, 4), "toeplitz_apply_ms": round(t_apply, 3),
            "toeplitz_passes": passes,
            "cg_plain_iterations": s["iterations"], "cg_plain_converged": s["converged"],
            "cg_plain_s": round(t_cg, 4), "cg_plain_rel_residual": s["rel_residual"],
            "cg_plain_true_rel_residual": s["true_rel_residual"],
            "time_to_1e-5_plain_s": round(ms * 1e-3 + t_q + t_plan + t_cg, 4),
            "cg_chan_iterations": s2["iterations"], "cg_chan_s": round(t_pcg, 4),
            "cg_chan_true_rel_residual": s2["true_rel_residual"],
            "time_to_1e-5_chan_s": round(ms * 1e-3 + t_q + t_plan + t_pcg, 4),
            "learned": learned,
            "gridding": gridding,
        }
        # F^H d needs the data, Q_l only the trajectory: a caller can run them
        # on two contexts (streams) at once; measured here from enqueue to
        # both done
        ctx2 = impatient.Context(0)
        dqr2, dqi2 = ctx2.alloc(((L + 1) * nq,)), ctx2.alloc(((L + 1) * nq,))

        def q_call2():
            call("imp_toeplitz_q", ctx2.h, C.c_int(nk), ptr(dkx), ptr(dky), None, ptr(dt_), None,
                 None, C.c_int(prob.Nx), C.c_int(prob.Ny), C.c_int(1), C.c_int(nq), None, None,
                 None, C.c_int(L), C.c_int(0), ptr(dqr2), ptr(dqi2))
        q_call2()
        ctx2.synchronize()
        t_both = []
        for _ in range(2):
            ctx.synchronize()
            t0 = time.perf_counter()
            fhd()
            q_call2()
            ctx.synchronize()
            ctx2.synchronize()
            t_both.append(time.perf_counter() - t0)
        t_both = min(t_both)
        ctx2.close()
        result["recon"]["fhd_and_q_two_streams_s"] = round(t_both, 4)
        log(f"F^H d and Q_l on two streams: {t_both * 1e3:.1f} ms (in sequence "
            f"{(ms * 1e-3 + t_q) * 1e3:.1f} ms)")
        if isinstance(learned, dict) and "time_to_1e-5_s" in learned:
            result["recon"]["time_to_1e-5_learned_two_streams_s"] = round(
                t_both + t_plan + learned.get("precond_network_s", 0.0) + learned["solve_s"], 4)
            result["recon"]["speedup_learned_vs_plain_cg"] = round(t_cg / learned["solve_s"], 2)
            result["recon"]["speedup_learned_vs_plain_time_to_1e-5"] = round(
                result["recon"]["time_to_1e-5_plain_s"] / learned["time_to_1e-5_s"], 2)
            result["recon"]["speedup_learned_vs_chan_cg"] = round(t_pcg / learned["solve_s"], 2)
            result["recon"]["speedup_learned_vs_chan_time_to_1e-5"] = round(
                result["recon"]["time_to_1e-5_chan_s"] / learned["time_to_1e-5_s"], 2)

    if not args.no_recon and not args.no_batch:
        result["batch"] = batch_leg(ctx, log)

    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cb = cpu_baseline(prob, d)
        if cb is not None:
            result["cpu_baseline"] = {k: (round(v, 4) if isinstance(v, float) else v)
                                      for k, v in cb.items()}
            log(f"cpu baseline: {cb['value']:.3f} GSample*vox/s ({cb['sample']})")
            bf = cpu_bruteforce_cfg0(ctx)
            if bf is not None:
                result["cpu_baseline"]["bruteforce_small_case"] = bf
            vo = cpu_vector_ops(prob, fd_gpu_ms if not args.no_recon else None)
            if vo is not None:
                result["cpu_baseline"]["vector_ops_512x512"] = vo
                log(f"brute force, small case: CPU {bf['cpu_s']:.3f} s, GPU {bf['gpu_s']:.4f} s")
    print(json.dumps(result), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
