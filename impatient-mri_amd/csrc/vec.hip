// Vector kernels, the periodic finite-difference regularizer and the
// device-resident CG of the Toeplitz solver.
//
// The reference reduces dot products with a multi-pass in-place tree whose
// blocks read and write the same buffer (recon.cu:585-700,
// dotProductGpu.cu:200-278) and finishes the sum in FP32 on the host.  Here a
// dot product is one pass of per-workgroup FP64 partials (wavefront shuffles,
// then LDS) into a separate buffer and one single-workgroup FP64 finish; the
// CG scalars alpha, beta and the best-iterate test never leave the device.
#include "common.h"
#include "toeplitz.h"

#include <chrono>
#include <cmath>
#include <mutex>

namespace imp {

namespace {

constexpr int VT = 256;
constexpr int RED_BLOCKS = 1024;  // max partials per reduction

__global__ void k_split_to_c(const float *re, const float *im, float2 *out, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float2(re[i], im ? im[i] : 0.f);
}
__global__ void k_c_to_split(const float2 *in, float *re, float *im, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float2 v = in[i];
        re[i] = v.x;
        im[i] = v.y;
    }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// block-wide FP64 sum; result valid in thread 0
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double sm[VT / 64];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sm[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < VT / 64; ++w) s += sm[w];
    __syncthreads();
    return s;
}

// partials[block] = sum Re(a^H b)
__global__ __launch_bounds__(VT) void k_dot_re(const float2 *__restrict__ a,
                                               const float2 *__restrict__ b, long n,
                                               double *__restrict__ partials) {
    double s = 0.0;
    for (long i = (long)blockIdx.x * VT + threadIdx.x; i < n; i += (long)gridDim.x * VT) {
        const float2 x = a[i], y = b[i];
        s = fma((double)x.x, (double)y.x, s);
        s = fma((double)x.y, (double)y.y, s);
    }
    s = block_sum(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(VT) void k_sum_partials(const double *__restrict__ partials,
                                                     int np, double *__restrict__ out) {
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += VT) s += partials[i];
    s = block_sum(s);
    if (threadIdx.x == 0) *out = s;
}

__global__ void k_axpy_c(float alpha, const float2 *x, float2 *y, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float2 a = x[i];
        float2 b = y[i];
        b.x += alpha * a.x;
        b.y += alpha * a.y;
        y[i] = b;
    }
}

// s = lambda * sum_d D_d^T W D_d p (periodic), optionally added to out.
// (D_d p)(i) = p(i) - p(i+e_d);  (D_d^T t)(i) = t(i) - t(i-e_d)
__global__ void k_fd_reg(int N1, int N2, int N3, float lambda, const float2 *__restrict__ p,
                         const float *__restrict__ w, float2 *__restrict__ out, int accumulate) {
    const long n = (long)N1 * N2 * N3;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int z = (int)(i % N3);
    const int x = (int)((i / N3) % N2);
    const int y = (int)(i / ((long)N3 * N2));
    const float2 pc = p[i];
    const float wc = w ? w[i] : 1.f;
    float sx = 0.f, sy = 0.f;
    // axes: y (stride N2*N3), x (stride N3), z (stride 1, 3D only)
    const int ndim = N3 > 1 ? 3 : 2;
    for (int d = 0; d < ndim; ++d) {
        long ip, im;
        if (d == 0) {
            ip = (long)((y + 1) % N1) * N2 * N3 + (long)x * N3 + z;
            im = (long)((y + N1 - 1) % N1) * N2 * N3 + (long)x * N3 + z;
        } else if (d == 1) {
            ip = (long)y * N2 * N3 + (long)((x + 1) % N2) * N3 + z;
            im = (long)y * N2 * N3 + (long)((x + N2 - 1) % N2) * N3 + z;
        } else {
            ip = (long)y * N2 * N3 + (long)x * N3 + (z + 1) % N3;
            im = (long)y * N2 * N3 + (long)x * N3 + (z + N3 - 1) % N3;
        }
        const float2 pp = p[ip], pm = p[im];
        const float wm = w ? w[im] : 1.f;
        // w_i (p_i - p_{i+e}) - w_{i-e} (p_{i-e} - p_i)
        sx += wc * (pc.x - pp.x) - wm * (pm.x - pc.x);
        sy += wc * (pc.y - pp.y) - wm * (pm.y - pc.y);
    }
    float2 r = make_float2(lambda * sx, lambda * sy);
    if (accumulate) {
        const float2 o = out[i];
        r.x += o.x;
        r.y += o.y;
    }
    out[i] = r;
}

// w = 0.5 / sqrt(sum_d |D_d x|^2 + 1e-6)  (multiplyGpuKernel2d/3d)
__global__ void k_tv_weight(int N1, int N2, int N3, const float2 *__restrict__ xv,
                            float *__restrict__ w) {
    const long n = (long)N1 * N2 * N3;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int z = (int)(i % N3);
    const int x = (int)((i / N3) % N2);
    const int y = (int)(i / ((long)N3 * N2));
    const float2 c = xv[i];
    const long ih = (long)y * N2 * N3 + (long)((x + 1) % N2) * N3 + z;   // Dhori
    const long iv = (long)((y + 1) % N1) * N2 * N3 + (long)x * N3 + z;   // Dverti
    float2 dh = make_float2(c.x - xv[ih].x, c.y - xv[ih].y);
    float2 dv = make_float2(c.x - xv[iv].x, c.y - xv[iv].y);
    float s = dh.x * dh.x + dh.y * dh.y + dv.x * dv.x + dv.y * dv.y;
    if (N3 > 1) {
        const long iz = (long)y * N2 * N3 + (long)x * N3 + (z + 1) % N3;  // Dzaxis
        const float2 dz = make_float2(c.x - xv[iz].x, c.y - xv[iz].y);
        s += dz.x * dz.x + dz.y * dz.y;
    }
    w[i] = 0.5f / sqrtf(s + 0.000001f);
}

// ------------------------------------------------------------------ CG
// device scalar slots
enum {
    SC_BETA = 0,     // |r|^2 of the current residual
    SC_BNORM = 1,    // |b|^2
    SC_ALPHA = 2,
    SC_RATIO = 3,    // beta2 / beta
    SC_BEST = 4,     // best |r|^2 so far
    SC_IMPROVED = 5, // 1 if this iteration improved the best residual
    SC_BREAK = 6,    // breakdown flag
    SC_PQ = 7,
    SC_RHO = 8,      // r^H z (preconditioned) or |r|^2 (plain)
    SC_N = 9
};

__global__ void k_cg_init(double *sc, const double *partials, int np, int carry) {
    // r = b (x0 = 0) or r = b - A x0: partials hold |r|^2; bnorm set separately.
    // carry: keep beta, rho and the best residual of the previous solve on
    // this plan (the restarts of recon.cu:2043-2048 set them only when
    // restart == 0)
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < np; ++i) s += partials[i];
        if (!carry) {
            sc[SC_BETA] = s;
            sc[SC_BEST] = s;
            sc[SC_RHO] = s;
        }
        sc[SC_IMPROVED] = 0.0;
        sc[SC_BREAK] = 0.0;
    }
}

// PCG: rho = Re(r^H z) from partials; ratio = rho_new / rho_old (init: just rho)
__global__ __launch_bounds__(VT) void k_pcg_rho(const double *__restrict__ partials, int np,
                                                double *__restrict__ sc, int init) {
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += VT) s += partials[i];
    s = block_sum(s);
    if (threadIdx.x == 0) {
        if (!init) {
            const double rho = sc[SC_RHO];
            sc[SC_RATIO] = rho > 0.0 ? s / rho : 0.0;
        }
        sc[SC_RHO] = s;
        if (!(s > 0.0) && sc[SC_BETA] > 0.0) sc[SC_BREAK] = 1.0;  // M not positive definite
    }
}

__global__ __launch_bounds__(VT) void k_cg_alpha(const double *__restrict__ partials, int np,
                                                 double *__restrict__ sc) {
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += VT) s += partials[i];
    s = block_sum(s);
    if (threadIdx.x == 0) {
        sc[SC_PQ] = s;
        // As the reference (recon.cu:2071-2077), alpha = rho / p^H A p of either
        // sign: the operator is positive semidefinite in exact arithmetic, but
        // its smallest eigenvalues (unsampled k-space, lambda-sized) sit below
        // the FP32 rounding of Q_l, so p^H A p < 0 can occur at tight
        // tolerances; the best-iterate tracking keeps the result.  Only a zero
        // or non-finite p^H A p stops the iteration.
        if (s == 0.0 || !isfinite(s)) {
            sc[SC_BREAK] = 1.0;
            sc[SC_ALPHA] = 0.0;
        } else {
            sc[SC_ALPHA] = sc[SC_RHO] / s;
        }
    }
}

// x += alpha p; r -= alpha q; partial |r|^2
__global__ __launch_bounds__(VT) void k_cg_update_xr(float2 *__restrict__ x, float2 *__restrict__ r,
                                                     const float2 *__restrict__ p,
                                                     const float2 *__restrict__ q, long n,
                                                     const double *__restrict__ sc,
                                                     double *__restrict__ partials) {
    const float alpha = (float)sc[SC_ALPHA];
    double s = 0.0;
    for (long i = (long)blockIdx.x * VT + threadIdx.x; i < n; i += (long)gridDim.x * VT) {
        const float2 pv = p[i], qv = q[i];
        float2 xv = x[i], rv = r[i];
        xv.x += alpha * pv.x;
        xv.y += alpha * pv.y;
        rv.x -= alpha * qv.x;
        rv.y -= alpha * qv.y;
        x[i] = xv;
        r[i] = rv;
        s = fma((double)rv.x, (double)rv.x, s);
        s = fma((double)rv.y, (double)rv.y, s);
    }
    s = block_sum(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// new |r|^2: best-iterate tracking; plain CG (pcg = 0) also sets
// ratio = |r_new|^2 / |r_old|^2 and rho
__global__ __launch_bounds__(VT) void k_cg_beta(const double *__restrict__ partials, int np,
                                                double *__restrict__ sc, int pcg) {
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += VT) s += partials[i];
    s = block_sum(s);
    if (threadIdx.x == 0) {
        if (!pcg) {
            const double beta = sc[SC_BETA];
            sc[SC_RATIO] = beta > 0.0 ? s / beta : 0.0;
            sc[SC_RHO] = s;
        }
        sc[SC_BETA] = s;
        if (s < sc[SC_BEST]) {
            sc[SC_BEST] = s;
            sc[SC_IMPROVED] = 1.0;
        } else {
            sc[SC_IMPROVED] = 0.0;
        }
    }
}

// p = r + ratio p (r: the residual, or M^-1 r in PCG); if improved: best = x
__global__ void k_cg_update_p(float2 *__restrict__ p, const float2 *__restrict__ r,
                              const float2 *__restrict__ x, float2 *__restrict__ best, long n,
                              const double *__restrict__ sc, int track_best) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float ratio = (float)sc[SC_RATIO];
    const float2 rv = r[i];
    float2 pv = p[i];
    pv.x = rv.x + ratio * pv.x;
    pv.y = rv.y + ratio * pv.y;
    p[i] = pv;
    if (track_best && sc[SC_IMPROVED] != 0.0) best[i] = x[i];
}

__global__ void k_scale_by_w(float2 *__restrict__ z, const float *__restrict__ w, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float s = w[i];
        z[i] = make_float2(z[i].x * s, z[i].y * s);
    }
}

// T. Chan's optimal circulant of a Toeplitz matrix with kernel t (per axis
// c_j = ((N - j) t_j + j t_{j-N}) / N), applied on every axis to
// t = sum_l Q_l on the standard Q grid ([2N1][2N2][M3], offset -N per axis).
__global__ void k_chan_circulant(const float *__restrict__ q_r, const float *__restrict__ q_i,
                                 int nseg, int N1, int N2, int N3, float2 *__restrict__ c) {
    const long n = (long)N1 * N2 * N3;
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int M2 = 2 * N2, M3 = N3 > 1 ? 2 * N3 : 1;
    const long G = 2L * N1 * M2 * M3;
    const int j3 = (int)(j % N3), j2 = (int)((j / N3) % N2), j1 = (int)(j / ((long)N3 * N2));
    double ar = 0.0, ai = 0.0;
    for (int a1 = 0; a1 < 2; ++a1) {
        const double w1 = a1 ? (double)j1 / N1 : (double)(N1 - j1) / N1;
        if (w1 == 0.0) continue;
        const int m1 = j1 + (1 - a1) * N1;
        for (int a2 = 0; a2 < 2; ++a2) {
            const double w2 = a2 ? (double)j2 / N2 : (double)(N2 - j2) / N2;
            if (w2 == 0.0) continue;
            const int m2 = j2 + (1 - a2) * N2;
            for (int a3 = 0; a3 < (N3 > 1 ? 2 : 1); ++a3) {
                const double w3 = N3 > 1 ? (a3 ? (double)j3 / N3 : (double)(N3 - j3) / N3) : 1.0;
                if (w3 == 0.0) continue;
                const int m3 = N3 > 1 ? j3 + (1 - a3) * N3 : 0;
                const long idx = ((long)m1 * M2 + m2) * M3 + m3;
                double sr = 0.0, si = 0.0;
                for (int l = 0; l < nseg; ++l) {
                    sr += q_r[(long)l * G + idx];
                    si += q_i[(long)l * G + idx];
                }
                ar += w1 * w2 * w3 * sr;
                ai += w1 * w2 * w3 * si;
            }
        }
    }
    c[j] = make_float2((float)ar, (float)ai);
}

// pw = 1 / (nvox * max(Re(DFT c) + lambda * periodic-Laplacian eigenvalue,
// 1e-3 * DC eigenvalue)); the floor keeps M positive definite.
__global__ void k_precond_weights(const float2 *__restrict__ chat, int N1, int N2, int N3,
                                  float lambda, float *__restrict__ w) {
    const long n = (long)N1 * N2 * N3;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k3 = (int)(i % N3), k2 = (int)((i / N3) % N2), k1 = (int)(i / ((long)N3 * N2));
    double lap = 2.0 - 2.0 * cospi(2.0 * k1 / N1) + 2.0 - 2.0 * cospi(2.0 * k2 / N2);
    if (N3 > 1) lap += 2.0 - 2.0 * cospi(2.0 * k3 / N3);
    const double dc = (double)chat[0].x;
    const double eig = fmax((double)chat[i].x + (double)lambda * lap, 1e-3 * fabs(dc));
    w[i] = eig > 0.0 ? (float)(1.0 / (eig * (double)n)) : 1.f;
}

__global__ void k_sub_c(const float2 *a, const float2 *b, float2 *out, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float2(a[i].x - b[i].x, a[i].y - b[i].y);
}

// owns an instantiated CG iteration graph
// Stream capture on one host thread is invalidated by an allocation, a free
// or a synchronous copy made by another thread while it records.  Solves run
// from several threads (one context each) take this lock around capture,
// instantiation and destruction, so graphs are never recorded concurrently;
// callers that solve from several threads allocate before they start.
std::mutex g_capture_mu;

struct CgGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~CgGraph() {
        if (!exec && !graph) return;
        std::lock_guard<std::mutex> lk(g_capture_mu);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

__global__ void k_copy_c(const float2 *__restrict__ a, float2 *__restrict__ b, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

int red_blocks(long n) {
    long nb = (n + VT * 4 - 1) / (VT * 4);
    if (nb < 1) nb = 1;
    if (nb > RED_BLOCKS) nb = RED_BLOCKS;
    return (int)nb;
}

}  // namespace

int split_to_c(imp_ctx *ctx, const float *re, const float *im, float2 *out, long n) {
    hipLaunchKernelGGL(k_split_to_c, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, re, im,
                       out, n);
    IMP_CHECK_LAUNCH();
    return IMP_OK;
}

int c_to_split(imp_ctx *ctx, const float2 *in, float *re, float *im, long n) {
    hipLaunchKernelGGL(k_c_to_split, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, in, re,
                       im, n);
    IMP_CHECK_LAUNCH();
    return IMP_OK;
}

int fd_reg(imp_ctx *ctx, int N1, int N2, int N3, float lambda, const float2 *p, const float *w,
           float2 *out, bool accumulate) {
    const long n = (long)N1 * N2 * N3;
    const int rec = prof_begin(ctx, "k_fd_reg");
    hipLaunchKernelGGL(k_fd_reg, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, N1, N2, N3,
                       lambda, p, w, out, accumulate ? 1 : 0);
    IMP_CHECK_LAUNCH();
    prof_end(ctx, rec);
    return IMP_OK;
}

int tv_weight(imp_ctx *ctx, int N1, int N2, int N3, const float2 *x, float *w) {
    const long n = (long)N1 * N2 * N3;
    hipLaunchKernelGGL(k_tv_weight, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, N1, N2,
                       N3, x, w);
    IMP_CHECK_LAUNCH();
    return IMP_OK;
}

int dot_re(imp_ctx *ctx, const float2 *a, const float2 *b, long n, double *out) {
    IMP_TRY(ensure_dscratch(ctx, RED_BLOCKS + 16));
    const int nb = red_blocks(n);
    hipLaunchKernelGGL(k_dot_re, dim3(nb), dim3(VT), 0, ctx->stream, a, b, n, ctx->dscratch);
    IMP_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(VT), 0, ctx->stream, ctx->dscratch, nb,
                       ctx->dscratch + RED_BLOCKS);
    IMP_CHECK_LAUNCH();
    IMP_HIP(hipMemcpyAsync(ctx->host_scalars, ctx->dscratch + RED_BLOCKS, sizeof(double),
                           hipMemcpyDeviceToHost, ctx->stream));
    IMP_HIP(hipStreamSynchronize(ctx->stream));
    *out = ctx->host_scalars[0];
    return IMP_OK;
}

int axpy_c(imp_ctx *ctx, float alpha, const float2 *x, float2 *y, long n) {
    hipLaunchKernelGGL(k_axpy_c, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, alpha, x, y,
                       n);
    IMP_CHECK_LAUNCH();
    return IMP_OK;
}

// out = A p = normal(p) + lambda R p
int toep_apply_full(imp_ctx *ctx, ToepPlan &P, const float2 *p, float2 *out) {
    if (P.sreg_lost)
        return fail(IMP_EINVAL, "Toeplitz operator: its sparse regularizer was destroyed");
    IMP_TRY(toep_apply_normal(ctx, P, p, out));
    if (P.lambda != 0.f) {
        if (P.sreg)
            IMP_TRY(sparse_reg(ctx, *P.sreg, P.lambda, p, P.w, out, true));
        else
            IMP_TRY(fd_reg(ctx, P.N1, P.N2, P.N3, P.lambda, p, P.w, out, true));
    }
    return IMP_OK;
}

int precond_chan(imp_ctx *ctx, ToepPlan &P, const float *q_r, const float *q_i) {
    const long n = P.nvox;
    float2 *c = P.vz;  // scratch: the circulant's first column, then its DFT
    hipLaunchKernelGGL(k_chan_circulant, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, q_r,
                       q_i, P.nseg, P.N1, P.N2, P.N3, c);
    IMP_CHECK_LAUNCH();
    IMP_TRY(fft_grid(ctx, c, P.N1, P.N2, P.N3, false));
    hipLaunchKernelGGL(k_precond_weights, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, c,
                       P.N1, P.N2, P.N3, P.lambda, P.pw);
    IMP_CHECK_LAUNCH();
    return IMP_OK;
}

int precond_apply(imp_ctx *ctx, ToepPlan &P, const float2 *r, float2 *z) {
    const long n = P.nvox;
    const int rec = prof_begin(ctx, "precond_apply");
    if (z != r) {
        hipLaunchKernelGGL(k_copy_c, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, r, z, n);
        IMP_CHECK_LAUNCH();
    }
    IMP_TRY(fft_grid(ctx, z, P.N1, P.N2, P.N3, false));
    hipLaunchKernelGGL(k_scale_by_w, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, z, P.pw,
                       n);
    IMP_CHECK_LAUNCH();
    IMP_TRY(fft_grid(ctx, z, P.N1, P.N2, P.N3, true));
    prof_end(ctx, rec);
    return IMP_OK;
}

int toep_cg(imp_ctx *ctx, ToepPlan &P, const float2 *b, float2 *x, int max_iter, double tol,
            int flags, imp_cg_stats *stats) {
    const auto t_start = std::chrono::steady_clock::now();
    const long n = P.nvox;
    const int nb = red_blocks(n);
    IMP_TRY(ensure_dscratch(ctx, RED_BLOCKS + 16));
    double *partials = ctx->dscratch;
    double *sc = P.scal;
    const bool best = (flags & IMP_CG_BEST_ITERATE) != 0;
    const bool pc = (flags & IMP_CG_PRECOND) != 0;
    if (pc && (!P.pw || !P.vz)) return fail(IMP_EINVAL, "PCG: operator has no preconditioner");
    const dim3 g1(ceil_div(n, 256)), b1(256);

    // |b|^2
    double bnorm = 0.0;
    IMP_TRY(dot_re(ctx, b, b, n, &bnorm));
    if (stats) *stats = imp_cg_stats{};
    if (!(bnorm > 0.0)) {
        IMP_HIP(hipMemsetAsync(x, 0, n * sizeof(float2), ctx->stream));
        IMP_HIP(hipStreamSynchronize(ctx->stream));
        if (stats) stats->converged = 1;
        return IMP_OK;
    }
    // r = b - A x0 (x0 = 0: r = b, the reference's wasted apply at
    // recon.cu:2036 is skipped since A 0 = 0)
    bool from_x0 = (flags & IMP_CG_USE_X0) != 0;
    double start_res = -1.0;
    if (from_x0) {
        IMP_TRY(toep_apply_full(ctx, P, x, P.vq));
        hipLaunchKernelGGL(k_sub_c, g1, b1, 0, ctx->stream, b, P.vq, P.vr, n);
        IMP_CHECK_LAUNCH();
        if (flags & IMP_CG_FALLBACK_ZERO) {
            // the zero start has residual |b|: keep x0 only if it does better
            double r0 = 0.0;
            IMP_TRY(dot_re(ctx, P.vr, P.vr, n, &r0));
            start_res = std::sqrt(r0 / bnorm);
            if (!(r0 < bnorm)) from_x0 = false;
        }
    }
    if (!from_x0) {
        IMP_HIP(hipMemsetAsync(x, 0, n * sizeof(float2), ctx->stream));
        IMP_HIP(hipMemcpyAsync(P.vr, b, n * sizeof(float2), hipMemcpyDeviceToDevice, ctx->stream));
    }
    IMP_CHECK_LAUNCH();
    IMP_HIP(hipMemcpyAsync(P.vbest, x, n * sizeof(float2), hipMemcpyDeviceToDevice, ctx->stream));
    hipLaunchKernelGGL(k_dot_re, dim3(nb), dim3(VT), 0, ctx->stream, P.vr, P.vr, n, partials);
    IMP_CHECK_LAUNCH();
    const bool carry = (flags & CG_CARRY_SCALARS) != 0 && !pc;
    hipLaunchKernelGGL(k_cg_init, dim3(1), dim3(1), 0, ctx->stream, sc, partials, nb,
                       carry ? 1 : 0);
    IMP_CHECK_LAUNCH();
    if (pc) {  // z = M^-1 r, p = z, rho = Re(r^H z)
        IMP_TRY(precond_apply(ctx, P, P.vr, P.vz));
        hipLaunchKernelGGL(k_dot_re, dim3(nb), dim3(VT), 0, ctx->stream, P.vr, P.vz, n, partials);
        hipLaunchKernelGGL(k_pcg_rho, dim3(1), dim3(VT), 0, ctx->stream, partials, nb, sc, 1);
        IMP_CHECK_LAUNCH();
        IMP_HIP(hipMemcpyAsync(P.vp, P.vz, n * sizeof(float2), hipMemcpyDeviceToDevice,
                               ctx->stream));
    } else {
        IMP_HIP(hipMemcpyAsync(P.vp, P.vr, n * sizeof(float2), hipMemcpyDeviceToDevice,
                               ctx->stream));
    }

    // one iteration: every launch is a kernel on ctx->stream with device
    // scalars, so the sequence can be captured into a HIP graph
    auto iteration = [&]() -> int {
        IMP_TRY(toep_apply_full(ctx, P, P.vp, P.vq));
        hipLaunchKernelGGL(k_dot_re, dim3(nb), dim3(VT), 0, ctx->stream, P.vp, P.vq, n, partials);
        hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(VT), 0, ctx->stream, partials, nb, sc);
        hipLaunchKernelGGL(k_cg_update_xr, dim3(nb), dim3(VT), 0, ctx->stream, x, P.vr, P.vp,
                           P.vq, n, sc, partials);
        hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(VT), 0, ctx->stream, partials, nb, sc,
                           pc ? 1 : 0);
        if (pc) {
            IMP_TRY(precond_apply(ctx, P, P.vr, P.vz));
            hipLaunchKernelGGL(k_dot_re, dim3(nb), dim3(VT), 0, ctx->stream, P.vr, P.vz, n,
                               partials);
            hipLaunchKernelGGL(k_pcg_rho, dim3(1), dim3(VT), 0, ctx->stream, partials, nb, sc, 0);
        }
        hipLaunchKernelGGL(k_cg_update_p, g1, b1, 0, ctx->stream, P.vp, pc ? P.vz : P.vr, x,
                           P.vbest, n, sc, best ? 1 : 0);
        IMP_CHECK_LAUNCH();
        return IMP_OK;
    };
    CgGraph graph;
    if (flags & IMP_CG_GRAPH) {
        // kernel timer events are not recorded into the graph
        const bool prof = ctx->prof_on;
        ctx->prof_on = false;
        std::lock_guard<std::mutex> lk(g_capture_mu);
        IMP_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const int st_cap = iteration();
        hipGraph_t g = nullptr;
        const hipError_t e_end = hipStreamEndCapture(ctx->stream, &g);
        ctx->prof_on = prof;
        graph.graph = g;
        if (st_cap != IMP_OK) return st_cap;
        if (e_end != hipSuccess) return fail(IMP_EHIP, "CG graph capture: %s", hipGetErrorString(e_end));
        IMP_HIP(hipGraphInstantiate(&graph.exec, graph.graph, nullptr, nullptr, 0));
    }

    int iter = 0, best_iter = 0;
    bool converged = false;
    double beta2 = 0.0;
    while (iter < max_iter) {
        if (graph.exec) {
            IMP_HIP(hipGraphLaunch(graph.exec, ctx->stream));
        } else {
            IMP_TRY(iteration());
        }
        IMP_HIP(hipMemcpyAsync(ctx->host_scalars, sc, SC_N * sizeof(double),
                               hipMemcpyDeviceToHost, ctx->stream));
        IMP_HIP(hipStreamSynchronize(ctx->stream));
        ++iter;
        const double *h = ctx->host_scalars;
        if (h[SC_BREAK] != 0.0)
            return fail(IMP_EBREAKDOWN, "CG breakdown at iteration %d: p^H A p = %g, r^H z = %g",
                        iter, h[SC_PQ], h[SC_RHO]);
        beta2 = h[SC_BETA];
        if (h[SC_IMPROVED] != 0.0) best_iter = iter;
        if (std::sqrt(beta2 / bnorm) < tol) {
            converged = true;
            break;
        }
    }
    const double best_res = ctx->host_scalars[SC_BEST];
    if (best && !(flags & CG_KEEP_LAST)) IMP_HIP(hipMemcpyAsync(x, P.vbest, n * sizeof(float2), hipMemcpyDeviceToDevice,
                                     ctx->stream));
    if (stats) {
        stats->iterations = iter;
        stats->converged = converged ? 1 : 0;
        stats->best_iteration = best ? best_iter : iter;
        stats->rel_residual = std::sqrt(beta2 / bnorm);
        stats->best_rel_residual = std::sqrt(best_res / bnorm);
        stats->true_rel_residual = -1.0;
        stats->start_used = from_x0 ? 1 : 0;
        stats->start_rel_residual = start_res;
    }
    if (flags & IMP_CG_TRUE_RESIDUAL) {
        IMP_TRY(toep_apply_full(ctx, P, x, P.vq));
        hipLaunchKernelGGL(k_sub_c, g1, b1, 0, ctx->stream, P.vq, b, P.vr, n);
        IMP_CHECK_LAUNCH();
        double rr = 0.0;
        IMP_TRY(dot_re(ctx, P.vr, P.vr, n, &rr));
        if (stats) stats->true_rel_residual = std::sqrt(rr / bnorm);
    }
    IMP_HIP(hipStreamSynchronize(ctx->stream));
    if (stats)
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start)
                             .count();
    return IMP_OK;
}

}  // namespace imp
