// Brute-force CG: bruteForceGpu's conjugate-gradient loop
// (bruteForce/bruteForceGpu.cu:470-654) on the direct NUDFT operators.
// Complex vectors are kept as [re | im] halves of one buffer so the
// real-arithmetic kernels (dot, axpby) see one array of 2n floats; dot
// products are FP64.
#include "common.h"
#include "toeplitz.h"

#include <cmath>
#include <vector>

namespace imp {

namespace {

constexpr int BT = 256;

// z = a*x + b*y over n floats
__global__ void k_axpby(float a, const float *x, float b, const float *y, float *z, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = a * x[i] + b * y[i];
}

__global__ __launch_bounds__(BT) void k_dot_real(const float *__restrict__ a,
                                                 const float *__restrict__ b, long n,
                                                 double *__restrict__ partials) {
    double s = 0.0;
    for (long i = (long)blockIdx.x * BT + threadIdx.x; i < n; i += (long)gridDim.x * BT)
        s = fma((double)a[i], (double)b[i], s);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    __shared__ double sm[BT / 64];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < BT / 64; ++w) t += sm[w];
        partials[blockIdx.x] = t;
    }
}

__global__ void k_finish(const double *partials, int np, double *out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < np; ++i) s += partials[i];
        *out = s;
    }
}

__global__ void k_split_pair_to_c(const float *v, long n, float2 *out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float2(v[i], v[n + i]);
}
__global__ void k_c_to_split_pair(const float2 *in, long n, float *v) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        v[i] = in[i].x;
        v[n + i] = in[i].y;
    }
}

struct BF {
    imp_ctx *ctx;
    int nb_max = 512;
    int axpby(float a, const float *x, float b, const float *y, float *z, long n) {
        hipLaunchKernelGGL(k_axpby, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, a, x, b,
                           y, z, n);
        IMP_CHECK_LAUNCH();
        return IMP_OK;
    }
    int dot(const float *a, const float *b, long n, double *out) {
        IMP_TRY(ensure_dscratch(ctx, 1024 + 16));
        long nb = (n + BT * 4 - 1) / (BT * 4);
        if (nb > nb_max) nb = nb_max;
        if (nb < 1) nb = 1;
        hipLaunchKernelGGL(k_dot_real, dim3((int)nb), dim3(BT), 0, ctx->stream, a, b, n,
                           ctx->dscratch);
        hipLaunchKernelGGL(k_finish, dim3(1), dim3(1), 0, ctx->stream, ctx->dscratch, (int)nb,
                           ctx->dscratch + 1024);
        IMP_CHECK_LAUNCH();
        IMP_HIP(hipMemcpyAsync(ctx->host_scalars, ctx->dscratch + 1024, sizeof(double),
                               hipMemcpyDeviceToHost, ctx->stream));
        IMP_HIP(hipStreamSynchronize(ctx->stream));
        *out = ctx->host_scalars[0];
        return IMP_OK;
    }
};

}  // namespace

int bruteforce_cg(imp_ctx *ctx, int num_k, int num_x, int ncoils, const float *kx,
                  const float *ky, const float *t, const float *x, const float *y,
                  const float *fm, const float *d_r, const float *d_i, const float *s_r,
                  const float *s_i, int cg_num, float lambda, const SparseReg *sreg,
                  int tv_num, float *x_r, float *x_i, imp_tv_pass_hook hook, void *hook_user) {
    if (!ctx || num_k <= 0 || num_x <= 0 || ncoils <= 0 || !kx || !ky || !t || !x || !y ||
        !d_r || !d_i || !x_r || !x_i || (s_r == nullptr) != (s_i == nullptr))
        return fail(IMP_EINVAL, "imp_bruteforce_cg: bad arguments");
    const int nimg = (int)std::lround(std::sqrt((double)num_x));
    const bool reg = lambda > 0.f;
    const int passes = tv_num > 1 ? tv_num : 1;
    if (reg && (!sreg || passes > 1) && (long)nimg * nimg != num_x)
        return fail(IMP_EINVAL, "imp_bruteforce_cg: regularizer needs a square image (%d)", num_x);
    // the sparse C of a 2D slice has two blocks of num_x rows (horizontal,
    // vertical differences; bruteForceGpu.cu:379-381, 513-541)
    if (sreg && (sreg->cols != num_x || sreg->rows != 2 * num_x))
        return fail(IMP_EINVAL, "imp_bruteforce_cg: C is %d x %d, expected %d x %d", sreg->rows,
                    sreg->cols, 2 * num_x, num_x);
    const long nk = (long)ncoils * num_k;
    const long nx = num_x;
    // buffers: y, Af, yAf, q  [2nk];  X, GtAf, grad, d, CtCf, CtCd [2nx];  cbuf [2nx float2];
    // TV weight [nx]
    const size_t bytes =
        (4 * 2 * nk + 6 * 2 * nx) * sizeof(float) + 2 * nx * sizeof(float2) + nx * sizeof(float);
    void *ws = nullptr;
    IMP_TRY(workspace(ctx, WS_BFORCE, bytes, &ws));
    float *yv = (float *)ws, *Af = yv + 2 * nk, *yAf = Af + 2 * nk, *q = yAf + 2 * nk;
    float *X = q + 2 * nk, *GtAf = X + 2 * nx, *grad = GtAf + 2 * nx, *dv = grad + 2 * nx,
          *CtCf = dv + 2 * nx, *CtCd = CtCf + 2 * nx;
    float2 *cb = (float2 *)(CtCd + 2 * nx), *cb2 = cb + nx;
    float *wtv = (float *)(cb2 + nx);
    const float *W = nullptr;  // W = 1 until a TV pass sets it
    hipStream_t st = ctx->stream;
    IMP_HIP(hipMemcpyAsync(yv, d_r, nk * sizeof(float), hipMemcpyDeviceToDevice, st));
    IMP_HIP(hipMemcpyAsync(yv + nk, d_i, nk * sizeof(float), hipMemcpyDeviceToDevice, st));
    IMP_HIP(hipMemcpyAsync(X, x_r, nx * sizeof(float), hipMemcpyDeviceToDevice, st));
    IMP_HIP(hipMemcpyAsync(X + nx, x_i, nx * sizeof(float), hipMemcpyDeviceToDevice, st));
    BF bf{ctx};

    auto F = [&](const float *img, float *kdat) {
        return nudft_forward(ctx, num_k, num_x, ncoils, kx, ky, nullptr, t, x, y, nullptr, fm,
                             img, img + nx, s_r, s_i, kdat, kdat + nk);
    };
    auto FH = [&](const float *kdat, float *img) {
        return nudft_adjoint(ctx, num_k, num_x, ncoils, kx, ky, nullptr, t, x, y, nullptr, fm,
                             nullptr, nullptr, kdat, kdat + nk, s_r, s_i, img, img + nx);
    };
    auto R = [&](const float *img, float *out) -> int {
        hipLaunchKernelGGL(k_split_pair_to_c, dim3(ceil_div(nx, 256)), dim3(256), 0, st, img,
                           nx, cb);
        IMP_CHECK_LAUNCH();
        if (sreg)
            IMP_TRY(sparse_reg(ctx, *sreg, lambda, cb, W, cb2, false));
        else
            IMP_TRY(fd_reg(ctx, nimg, nimg, 1, lambda, cb, W, cb2, false));
        hipLaunchKernelGGL(k_c_to_split_pair, dim3(ceil_div(nx, 256)), dim3(256), 0, st, cb2,
                           nx, out);
        IMP_CHECK_LAUNCH();
        return IMP_OK;
    };

    // TV passes (bruteForceGpu.cu:431-757): pass 0 with W = 1, each later
    // pass with W = 0.5 / sqrt(|D_h x|^2 + |D_v x|^2 + 1e-6) of the current
    // image, and the CG restarted from it; one pass without TV
    std::vector<float> hx;
    for (int tv_i = 0; tv_i < passes; ++tv_i) {
    if (hook) {
        // the image this pass starts from, handed to the caller on the host
        // (the -tv_update export of bruteForceGpu.cu:452-463)
        hx.resize(2 * nx);
        IMP_HIP(hipMemcpyAsync(hx.data(), X, 2 * nx * sizeof(float), hipMemcpyDeviceToHost, st));
        IMP_HIP(hipStreamSynchronize(st));
        hook(hook_user, tv_i, hx.data(), hx.data() + nx, num_x);
    }
    if (tv_i > 0 && reg) {
        hipLaunchKernelGGL(k_split_pair_to_c, dim3(ceil_div(nx, 256)), dim3(256), 0, st, X, nx,
                           cb);
        IMP_CHECK_LAUNCH();
        IMP_TRY(tv_weight(ctx, nimg, nimg, 1, cb, wtv));
        W = wtv;
    }
    IMP_TRY(F(X, Af));
    double oldinprod = 0.0;
    for (int cg_i = 1; cg_i < cg_num; ++cg_i) {
        IMP_TRY(bf.axpby(1.f, yv, -1.f, Af, yAf, 2 * nk));  // y - Af
        IMP_TRY(FH(yAf, GtAf));
        if (reg) {
            IMP_TRY(R(X, CtCf));
            IMP_TRY(bf.axpby(1.f, GtAf, -1.f, CtCf, grad, 2 * nx));
        } else {
            IMP_HIP(hipMemcpyAsync(grad, GtAf, 2 * nx * sizeof(float), hipMemcpyDeviceToDevice,
                                   st));
        }
        double newinprod = 0.0;
        IMP_TRY(bf.dot(grad, grad, 2 * nx, &newinprod));
        if (cg_i == 1) {
            IMP_HIP(hipMemcpyAsync(dv, grad, 2 * nx * sizeof(float), hipMemcpyDeviceToDevice, st));
        } else {
            const double gamma = oldinprod == 0.0 ? 0.0 : newinprod / oldinprod;
            IMP_TRY(bf.axpby(1.f, grad, (float)gamma, dv, dv, 2 * nx));
        }
        oldinprod = newinprod;
        IMP_TRY(F(dv, q));
        double a_ctcd = 0.0;
        if (reg) {
            IMP_TRY(R(dv, CtCd));
            IMP_TRY(bf.dot(CtCd, dv, 2 * nx, &a_ctcd));
        }
        double a_dgrad = 0.0, a_q = 0.0;
        IMP_TRY(bf.dot(dv, grad, 2 * nx, &a_dgrad));
        IMP_TRY(bf.dot(q, q, 2 * nk, &a_q));
        const double denom = a_q + a_ctcd;
        if (!(denom > 0.0) || !std::isfinite(denom)) {
            if (a_dgrad == 0.0) break;  // exact solution reached
            return fail(IMP_EBREAKDOWN, "brute-force CG breakdown at iteration %d", cg_i);
        }
        const float alpha = (float)(a_dgrad / denom);
        IMP_TRY(bf.axpby(1.f, Af, alpha, q, Af, 2 * nk));
        IMP_TRY(bf.axpby(1.f, X, alpha, dv, X, 2 * nx));
    }
    }
    IMP_HIP(hipMemcpyAsync(x_r, X, nx * sizeof(float), hipMemcpyDeviceToDevice, st));
    IMP_HIP(hipMemcpyAsync(x_i, X + nx, nx * sizeof(float), hipMemcpyDeviceToDevice, st));
    IMP_HIP(hipStreamSynchronize(st));
    return IMP_OK;
}

}  // namespace imp

extern "C" int imp_bruteforce_cg(imp_ctx *ctx, int num_k, int num_x, int ncoils,
                                 const float *kx, const float *ky, const float *t,
                                 const float *x, const float *y, const float *fm,
                                 const float *d_r, const float *d_i, const float *s_r,
                                 const float *s_i, int cg_num, float lambda, float *x_r,
                                 float *x_i) {
    return imp::bruteforce_cg(ctx, num_k, num_x, ncoils, kx, ky, t, x, y, fm, d_r, d_i, s_r,
                              s_i, cg_num, lambda, nullptr, 1, x_r, x_i, nullptr, nullptr);
}

extern "C" int imp_bruteforce_cg_ex(imp_ctx *ctx, int num_k, int num_x, int ncoils,
                                    const float *kx, const float *ky, const float *t,
                                    const float *x, const float *y, const float *fm,
                                    const float *d_r, const float *d_i, const float *s_r,
                                    const float *s_i, int cg_num, float lambda,
                                    const imp_sparse_reg *reg, int tv_num, float *x_r,
                                    float *x_i) {
    const imp::SparseReg *R = imp::sparse_reg_of(reg);
    if (reg && !R)
        return imp::fail(IMP_EINVAL, "imp_bruteforce_cg_ex: regularizer has no context");
    // the regularizer's arrays are used on ctx's stream: it must belong to ctx
    if (R && (!ctx || imp::sparse_reg_ctx(reg) != ctx))
        return imp::fail(IMP_EINVAL, "imp_bruteforce_cg_ex: regularizer belongs to another "
                                     "context");
    return imp::bruteforce_cg(ctx, num_k, num_x, ncoils, kx, ky, t, x, y, fm, d_r, d_i, s_r,
                              s_i, cg_num, lambda, R, tv_num, x_r, x_i, nullptr, nullptr);
}
