// Context, errors, workspace and the device-pointer C entry points.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "common.h"
#include "fftcore.h"
#include "toeplitz.h"

namespace imp {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }
const char *last_error() { return g_last_error.c_str(); }

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int workspace(imp_ctx *ctx, int slot, size_t bytes, void **out) {
    Workspace &w = ctx->ws[slot];
    if (w.bytes < bytes) {
        if (w.ptr) {
            IMP_HIP(hipStreamSynchronize(ctx->stream));
            IMP_HIP(hipFree(w.ptr));
            w.ptr = nullptr;
            w.bytes = 0;
        }
        const size_t want = bytes + bytes / 8;  // slack for slowly growing calls
        if (hipMalloc(&w.ptr, want) != hipSuccess) {
            (void)hipGetLastError();
            w.ptr = nullptr;
            return fail(IMP_ENOMEM, "workspace: cannot allocate %zu bytes", want);
        }
        w.bytes = want;
    }
    *out = w.ptr;
    return IMP_OK;
}

int workspace_release(imp_ctx *ctx, int slot, size_t keep) {
    Workspace &w = ctx->ws[slot];
    if (w.ptr && w.bytes > keep) {
        IMP_HIP(hipStreamSynchronize(ctx->stream));
        IMP_HIP(hipFree(w.ptr));
        w.ptr = nullptr;
        w.bytes = 0;
    }
    return IMP_OK;
}

int ensure_dscratch(imp_ctx *ctx, size_t n) {
    if (ctx->dscratch_n >= n) return IMP_OK;
    if (ctx->dscratch) {
        IMP_HIP(hipStreamSynchronize(ctx->stream));
        IMP_HIP(hipFree(ctx->dscratch));
    }
    ctx->dscratch = nullptr;
    IMP_HIP(hipMalloc(&ctx->dscratch, n * sizeof(double)));
    ctx->dscratch_n = n;
    return IMP_OK;
}

}  // namespace imp

using namespace imp;

// ------------------------------------------------------------------ basics
extern "C" const char *imp_version_string(void) { return "impatient-mi355x 0.1.0 (gfx950)"; }

extern "C" const char *imp_status_string(int s) {
    switch (s) {
        case IMP_OK: return "ok";
        case IMP_EINVAL: return "invalid argument";
        case IMP_EHIP: return "HIP runtime error";
        case IMP_EIO: return "I/O error";
        case IMP_ENOMEM: return "out of memory";
        case IMP_EFORMAT: return "malformed data file";
        case IMP_ENODEV: return "no usable GPU";
        case IMP_EBREAKDOWN: return "CG breakdown";
        case IMP_EUNSUPPORTED: return "unsupported size or option";
        default: return "unknown status";
    }
}

extern "C" const char *imp_last_error(void) { return imp::last_error(); }

namespace imp {

int prof_begin(imp_ctx *ctx, const char *name) {
    if (!ctx->prof_on) return -1;
    hipEvent_t ev[2];
    for (int k = 0; k < 2; ++k) {
        if (!ctx->prof_free.empty()) {
            ev[k] = ctx->prof_free.back();
            ctx->prof_free.pop_back();
        } else if (hipEventCreate(&ev[k]) != hipSuccess) {
            (void)hipGetLastError();
            if (k == 1) ctx->prof_free.push_back(ev[0]);
            return -1;  // timing is best effort; the launch itself proceeds
        }
    }
    (void)hipEventRecord(ev[0], ctx->stream);
    ctx->prof_recs.push_back(ProfRec{name, ev[0], ev[1]});
    return (int)ctx->prof_recs.size() - 1;
}

void prof_end(imp_ctx *ctx, int rec) {
    if (rec < 0 || rec >= (int)ctx->prof_recs.size()) return;
    (void)hipEventRecord(ctx->prof_recs[rec].end, ctx->stream);
}

static void prof_release(imp_ctx *ctx) {
    for (auto &r : ctx->prof_recs) {
        ctx->prof_free.push_back(r.begin);
        ctx->prof_free.push_back(r.end);
    }
    ctx->prof_recs.clear();
}

}  // namespace imp

extern "C" int imp_ctx_create(int device, imp_ctx **out) {
    if (!out) return fail(IMP_EINVAL, "imp_ctx_create: out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return fail(IMP_ENODEV, "imp_ctx_create: no HIP device");
    }
    if (device < 0 || device >= n)
        return fail(IMP_EINVAL, "imp_ctx_create: device %d out of range [0,%d)", device, n);
    IMP_HIP(hipSetDevice(device));
    imp_ctx *ctx = new (std::nothrow) imp_ctx();
    if (!ctx) return fail(IMP_ENOMEM, "imp_ctx_create: host allocation");
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void **)&ctx->host_scalars, 64 * sizeof(double), hipHostMallocDefault) !=
            hipSuccess) {
        (void)hipGetLastError();
        delete ctx;
        return fail(IMP_EHIP, "imp_ctx_create: stream / pinned allocation failed");
    }
    // reduction partials (up to 1024 blocks) with their sums, and the FFT
    // twiddles, made here so that solves never allocate on first use
    if (ensure_dscratch(ctx, 1024 + 16) != IMP_OK || twiddle_table_create(ctx) != IMP_OK) {
        (void)hipGetLastError();
        if (ctx->dscratch) (void)hipFree(ctx->dscratch);
        if (ctx->twiddle) (void)hipFree(ctx->twiddle);
        (void)hipStreamDestroy(ctx->stream);
        (void)hipHostFree(ctx->host_scalars);
        delete ctx;
        return fail(IMP_ENOMEM, "imp_ctx_create: reduction scratch / twiddle table");
    }
    *out = ctx;
    clear_error();
    return IMP_OK;
}

static void release_toep_memory(imp_toep *op) {
    for (void *p : op->owned) (void)hipFree(p);
    op->owned.clear();
}

extern "C" int imp_ctx_destroy(imp_ctx *ctx) {
    if (!ctx) return IMP_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    // operators still alive lose their device memory and their context;
    // a later imp_toep_destroy only frees the host object
    for (imp_toep *op : ctx->toeps) {
        release_toep_memory(op);
        op->ctx = nullptr;
    }
    ctx->toeps.clear();
    for (imp_cnn *net : ctx->cnns) detach_cnn(net);
    ctx->cnns.clear();
    for (imp_sparse_reg *reg : ctx->sregs) detach_sparse_reg(reg);
    ctx->sregs.clear();
    prof_release(ctx);
    for (hipEvent_t e : ctx->prof_free) (void)hipEventDestroy(e);
    for (auto &w : ctx->ws)
        if (w.ptr) (void)hipFree(w.ptr);
    if (ctx->dscratch) (void)hipFree(ctx->dscratch);
    if (ctx->twiddle) (void)hipFree(ctx->twiddle);
    if (ctx->host_scalars) (void)hipHostFree(ctx->host_scalars);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return IMP_OK;
}

extern "C" int imp_ctx_stream(imp_ctx *ctx, void **s) {
    if (!ctx || !s) return fail(IMP_EINVAL, "imp_ctx_stream: NULL argument");
    *s = (void *)ctx->stream;
    return IMP_OK;
}

extern "C" int imp_profile_enable(imp_ctx *ctx, int on) {
    if (!ctx) return fail(IMP_EINVAL, "imp_profile_enable: NULL ctx");
    ctx->prof_on = on != 0;
    return IMP_OK;
}

extern "C" int imp_profile_reset(imp_ctx *ctx) {
    if (!ctx) return fail(IMP_EINVAL, "imp_profile_reset: NULL ctx");
    IMP_HIP(hipStreamSynchronize(ctx->stream));
    prof_release(ctx);
    return IMP_OK;
}

extern "C" int imp_profile_read(imp_ctx *ctx, const char *kernel, double *total_ms,
                                int *launches) {
    if (!ctx || !kernel || !total_ms || !launches)
        return fail(IMP_EINVAL, "imp_profile_read: NULL argument");
    IMP_HIP(hipStreamSynchronize(ctx->stream));
    double sum = 0.0;
    int n = 0;
    for (const auto &r : ctx->prof_recs) {
        if (std::strcmp(r.name, kernel) != 0) continue;
        float ms = 0.f;
        IMP_HIP(hipEventElapsedTime(&ms, r.begin, r.end));
        sum += ms;
        ++n;
    }
    *total_ms = sum;
    *launches = n;
    return IMP_OK;
}

extern "C" int imp_ctx_synchronize(imp_ctx *ctx) {
    if (!ctx) return fail(IMP_EINVAL, "imp_ctx_synchronize: NULL ctx");
    IMP_HIP(hipStreamSynchronize(ctx->stream));
    return IMP_OK;
}

extern "C" int imp_device_alloc(imp_ctx *ctx, size_t bytes, void **dptr) {
    if (!ctx || !dptr) return fail(IMP_EINVAL, "imp_device_alloc: NULL argument");
    IMP_HIP(hipSetDevice(ctx->device));
    if (hipMalloc(dptr, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        return fail(IMP_ENOMEM, "imp_device_alloc: %zu bytes", bytes);
    }
    return IMP_OK;
}

extern "C" int imp_device_free(imp_ctx *ctx, void *dptr) {
    if (!ctx) return fail(IMP_EINVAL, "imp_device_free: NULL ctx");
    if (dptr) {
        IMP_HIP(hipStreamSynchronize(ctx->stream));
        IMP_HIP(hipFree(dptr));
    }
    return IMP_OK;
}

extern "C" int imp_host_alloc(imp_ctx *ctx, size_t bytes, void **hptr) {
    if (!ctx || !hptr) return fail(IMP_EINVAL, "imp_host_alloc: NULL argument");
    IMP_HIP(hipSetDevice(ctx->device));
    if (hipHostMalloc(hptr, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        *hptr = nullptr;
        return fail(IMP_ENOMEM, "imp_host_alloc: %zu bytes", bytes);
    }
    return IMP_OK;
}

extern "C" int imp_host_free(imp_ctx *ctx, void *hptr) {
    if (!ctx) return fail(IMP_EINVAL, "imp_host_free: NULL ctx");
    if (hptr) {
        // a copy from it may still be queued on the stream
        IMP_HIP(hipStreamSynchronize(ctx->stream));
        IMP_HIP(hipHostFree(hptr));
    }
    return IMP_OK;
}

extern "C" int imp_copy_h2d(imp_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx) return fail(IMP_EINVAL, "imp_copy_h2d: NULL ctx");
    IMP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    IMP_HIP(hipStreamSynchronize(ctx->stream));
    return IMP_OK;
}

extern "C" int imp_copy_d2h(imp_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx) return fail(IMP_EINVAL, "imp_copy_d2h: NULL ctx");
    IMP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    IMP_HIP(hipStreamSynchronize(ctx->stream));
    return IMP_OK;
}

extern "C" int imp_memset_d(imp_ctx *ctx, void *dst, int value, size_t bytes) {
    if (!ctx) return fail(IMP_EINVAL, "imp_memset_d: NULL ctx");
    IMP_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
    return IMP_OK;
}

// --------------------------------------------------------------- operators
extern "C" int imp_nudft_adjoint(imp_ctx *ctx, int num_k, int num_x, int ncoils,
                                 const float *kx, const float *ky, const float *kz,
                                 const float *t, const float *x, const float *y,
                                 const float *z, const float *fm, const float *phi_r,
                                 const float *phi_i, const float *d_r, const float *d_i,
                                 const float *s_r, const float *s_i, float *out_r,
                                 float *out_i) {
    return nudft_adjoint(ctx, num_k, num_x, ncoils, kx, ky, kz, t, x, y, z, fm, phi_r, phi_i,
                         d_r, d_i, s_r, s_i, out_r, out_i);
}

extern "C" int imp_nudft_forward(imp_ctx *ctx, int num_k, int num_x, int ncoils,
                                 const float *kx, const float *ky, const float *kz,
                                 const float *t, const float *x, const float *y,
                                 const float *z, const float *fm, const float *rho_r,
                                 const float *rho_i, const float *s_r, const float *s_i,
                                 float *d_r, float *d_i) {
    return nudft_forward(ctx, num_k, num_x, ncoils, kx, ky, kz, t, x, y, z, fm, rho_r, rho_i,
                         s_r, s_i, d_r, d_i);
}

extern "C" int imp_toeplitz_q(imp_ctx *ctx, int num_k, const float *kx, const float *ky,
                              const float *kz, const float *t, const float *phi_r,
                              const float *phi_i, int Nx, int Ny, int Nz, int num_q,
                              const float *gx, const float *gy, const float *gz,
                              int ntime_segs, int flags, float *q_r, float *q_i) {
    if (!ctx || !t || num_k <= 0) return fail(IMP_EINVAL, "imp_toeplitz_q: bad arguments");
    float tt[2];
    IMP_HIP(hipMemcpyAsync(&tt[0], t, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    IMP_HIP(hipMemcpyAsync(&tt[1], t + (num_k - 1), sizeof(float), hipMemcpyDeviceToHost,
                           ctx->stream));
    IMP_HIP(hipStreamSynchronize(ctx->stream));
    return toeplitz_q(ctx, num_k, kx, ky, kz, t, phi_r, phi_i, Nx, Ny, Nz, num_q, gx, gy, gz,
                      ntime_segs, flags, tt, q_r, q_i);
}

extern "C" int imp_fd_reg_apply(imp_ctx *ctx, int N1, int N2, int N3, float lambda,
                                const float *p_r, const float *p_i, const float *w,
                                float *s_r, float *s_i) {
    if (!ctx || N1 <= 0 || N2 <= 0 || N3 <= 0 || !p_r || !p_i || !s_r || !s_i)
        return fail(IMP_EINVAL, "imp_fd_reg_apply: bad arguments");
    const long n = (long)N1 * N2 * N3;
    void *buf = nullptr;
    IMP_TRY(workspace(ctx, WS_REG, 2 * n * sizeof(float2), &buf));
    float2 *pc = (float2 *)buf, *sc = pc + n;
    IMP_TRY(split_to_c(ctx, p_r, p_i, pc, n));
    IMP_TRY(fd_reg(ctx, N1, N2, N3, lambda, pc, w, sc, false));
    IMP_TRY(c_to_split(ctx, sc, s_r, s_i, n));
    return IMP_OK;
}

extern "C" int imp_tv_weight(imp_ctx *ctx, int N1, int N2, int N3, const float *x_r,
                             const float *x_i, float *w) {
    if (!ctx || N1 <= 0 || N2 <= 0 || N3 <= 0 || !x_r || !x_i || !w)
        return fail(IMP_EINVAL, "imp_tv_weight: bad arguments");
    const long n = (long)N1 * N2 * N3;
    void *buf = nullptr;
    IMP_TRY(workspace(ctx, WS_REG, n * sizeof(float2), &buf));
    IMP_TRY(split_to_c(ctx, x_r, x_i, (float2 *)buf, n));
    IMP_TRY(tv_weight(ctx, N1, N2, N3, (const float2 *)buf, w));
    return IMP_OK;
}

// -------------------------------------------------------- Toeplitz operator
namespace {
__global__ void k_fm_to_cycles(const float *fm, float *f, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i] = fm ? (float)((double)fm[i] * (1.0 / (2.0 * M_PI))) : 0.f;
}

int toep_alloc(imp_toep *op, size_t bytes, void **p) {
    if (hipMalloc(p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return fail(IMP_ENOMEM, "imp_toep_create: cannot allocate %zu bytes", bytes);
    }
    op->owned.push_back(*p);
    return IMP_OK;
}
}  // namespace

extern "C" int imp_toep_destroy(imp_toep *op) {
    if (!op) return IMP_OK;
    if (op->ctx) {
        imp_ctx *ctx = op->ctx;
        (void)hipStreamSynchronize(ctx->stream);
        release_toep_memory(op);
        ctx->toeps.erase(op);
    }
    delete op;
    return IMP_OK;
}

extern "C" int imp_toep_create(imp_ctx *ctx, int N1, int N2, int N3, int ncoils,
                               int ntime_segs, float tau, float t0, const float *q_r,
                               const float *q_i, const float *fm, const float *s_r,
                               const float *s_i, float lambda, imp_toep **out) {
    if (!ctx || !out || N1 <= 0 || N2 <= 0 || N3 <= 0 || ncoils <= 0 || ntime_segs < 0 ||
        !q_r || !q_i || (s_r == nullptr) != (s_i == nullptr))
        return fail(IMP_EINVAL, "imp_toep_create: bad arguments");
    if (!toep_sizes_supported(N1, N2, N3))
        return fail(IMP_EUNSUPPORTED,
                    "imp_toep_create: image (%d,%d,%d) needs lengths 2^k or 3 * 2^k whose "
                    "doubles are at most 2048",
                    N1, N2, N3);
    *out = nullptr;
    imp_toep *op = new (std::nothrow) imp_toep();
    if (!op) return fail(IMP_ENOMEM, "imp_toep_create: host allocation");
    op->ctx = ctx;
    ctx->toeps.insert(op);
    ToepPlan &P = op->P;
    P.N1 = N1;
    P.N2 = N2;
    P.N3 = N3;
    P.M1 = 2 * N1;
    P.M2 = 2 * N2;
    P.M3 = N3 > 1 ? 2 * N3 : 1;
    P.G = (long)P.M1 * P.M2 * P.M3;
    P.nvox = (long)N1 * N2 * N3;
    P.nseg = ntime_segs + 1;
    P.ncoils = ncoils;
    P.tau = tau;
    P.t0 = t0;
    P.lambda = lambda;
    // coil groups for pass E: enough workgroups to fill the GPU, 16 waves
    // per CU (pass E holds row_threads / (threads per line) image lines per
    // workgroup)
    {
        const int Ml = N3 > 1 ? P.M3 : P.M2;
        const int rt = fftc::row_threads(Ml);
        const int nl = rt / fftc::threads_per_line(Ml);
        const long wgs_per_group = (long)ceil_div((long)N1 * (N3 > 1 ? N2 : 1), nl);
        const long target = (1024L / rt) * ctx->num_cus;
        int groups = 1;
        while (groups < ncoils && wgs_per_group * groups < target) groups *= 2;
        if (groups > ncoils) groups = ncoils;
        P.coils_per_group = ceil_div(ncoils, groups);
        P.coils_per_group_min = P.coils_per_group;
    }
    const int ngroups = ceil_div(ncoils, P.coils_per_group);
    const long A = N3 > 1 ? P.M2 : 1, Ml = N3 > 1 ? P.M3 : P.M2;
    int st = IMP_OK;
    void *p = nullptr;
#define ALLOC(field, bytes)                                   \
    if (st == IMP_OK && (st = toep_alloc(op, (bytes), &p)) == IMP_OK) field = (decltype(field))p;
    ALLOC(P.qf, (size_t)P.nseg * P.G * sizeof(float2));
    ALLOC(P.qfr, (size_t)P.nseg * P.G * sizeof(float));
    ALLOC(P.g1, (size_t)ncoils * N1 * A * Ml * sizeof(float2));
    ALLOC(P.partial, (size_t)ngroups * P.nvox * sizeof(float2));
    ALLOC(P.fcyc, (size_t)P.nvox * sizeof(float));
    if (s_r) ALLOC(P.sens, (size_t)ncoils * P.nvox * sizeof(float2));
    ALLOC(P.vb, (size_t)P.nvox * sizeof(float2));
    ALLOC(P.vx, (size_t)P.nvox * sizeof(float2));
    ALLOC(P.vr, (size_t)P.nvox * sizeof(float2));
    ALLOC(P.vp, (size_t)P.nvox * sizeof(float2));
    ALLOC(P.vq, (size_t)P.nvox * sizeof(float2));
    ALLOC(P.vbest, (size_t)P.nvox * sizeof(float2));
    ALLOC(P.vz, (size_t)P.nvox * sizeof(float2));
    ALLOC(P.pw, (size_t)P.nvox * sizeof(float));
    ALLOC(P.pw_chan, (size_t)P.nvox * sizeof(float));
    ALLOC(P.scal, 16 * sizeof(double));
#undef ALLOC
    if (st != IMP_OK) {
        imp_toep_destroy(op);
        return st;
    }
    hipLaunchKernelGGL(k_fm_to_cycles, dim3(ceil_div(P.nvox, 256)), dim3(256), 0, ctx->stream,
                       fm, P.fcyc, P.nvox);
    if (s_r) st = split_to_c(ctx, s_r, s_i, P.sens, (long)ncoils * P.nvox);
    if (st == IMP_OK) st = toep_prepare_qf(ctx, P, q_r, q_i);
    if (st == IMP_OK) st = precond_chan(ctx, P, q_r, q_i);
    if (st == IMP_OK &&
        hipMemcpyAsync(P.pw_chan, P.pw, (size_t)P.nvox * sizeof(float), hipMemcpyDeviceToDevice,
                       ctx->stream) != hipSuccess)
        st = fail(IMP_EHIP, "imp_toep_create: %s", hipGetErrorString(hipGetLastError()));
    if (st == IMP_OK && hipStreamSynchronize(ctx->stream) != hipSuccess)
        st = fail(IMP_EHIP, "imp_toep_create: %s", hipGetErrorString(hipGetLastError()));
    if (st != IMP_OK) {
        imp_toep_destroy(op);
        return st;
    }
    *out = op;
    return IMP_OK;
}

namespace {
__global__ void k_check_weights(const float *w, long n, unsigned int *bad) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(w[i] > 0.f && w[i] < INFINITY)) atomicOr(bad, 1u);
}
}  // namespace

extern "C" int imp_toep_set_precond(imp_toep *op, const float *w) {
    if (!op || !op->ctx) return fail(IMP_EINVAL, "imp_toep_set_precond: bad operator");
    ToepPlan &P = op->P;
    imp_ctx *ctx = op->ctx;
    const size_t bytes = (size_t)P.nvox * sizeof(float);
    if (!w) {
        IMP_HIP(hipMemcpyAsync(P.pw, P.pw_chan, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        IMP_HIP(hipStreamSynchronize(ctx->stream));
        return IMP_OK;
    }
    // reject weights that would make M indefinite (checked before copying)
    IMP_TRY(ensure_dscratch(ctx, 16));
    unsigned int *bad = (unsigned int *)ctx->dscratch;
    IMP_HIP(hipMemsetAsync(bad, 0, sizeof(unsigned int), ctx->stream));
    hipLaunchKernelGGL(k_check_weights, dim3(ceil_div(P.nvox, 256)), dim3(256), 0, ctx->stream, w,
                       P.nvox, bad);
    IMP_CHECK_LAUNCH();
    unsigned int hbad = 0;
    IMP_HIP(hipMemcpyAsync(&hbad, bad, sizeof(unsigned int), hipMemcpyDeviceToHost, ctx->stream));
    IMP_HIP(hipStreamSynchronize(ctx->stream));
    if (hbad)
        return fail(IMP_EINVAL, "imp_toep_set_precond: weights must be positive and finite");
    IMP_HIP(hipMemcpyAsync(P.pw, w, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    IMP_HIP(hipStreamSynchronize(ctx->stream));
    return IMP_OK;
}

namespace {
__device__ __forceinline__ float signed_freq(int k, int n) {
    return (float)(k < (n + 1) / 2 ? k : k - n) / (float)n;
}
// radial bin of frequency i (natural DFT order), rho = |k/N| in [0, rmax]
__global__ void k_precond_radial(const float *chan, const float *gains, int nbins, int N1,
                                 int N2, int N3, float *w) {
    const long n = (long)N1 * N2 * N3;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k3 = (int)(i % N3), k2 = (int)((i / N3) % N2), k1 = (int)(i / ((long)N3 * N2));
    const float f1 = signed_freq(k1, N1), f2 = signed_freq(k2, N2);
    const float f3 = N3 > 1 ? signed_freq(k3, N3) : 0.f;
    const int dims = 2 + (N3 > 1 ? 1 : 0);
    const float rmax = 0.5f * sqrtf((float)dims);
    const float rho = sqrtf(f1 * f1 + f2 * f2 + f3 * f3);
    const int b = min(nbins - 1, (int)(rho / rmax * nbins));
    w[i] = chan[i] * gains[b];
}
}  // namespace

extern "C" int imp_toep_set_precond_radial(imp_toep *op, const float *gains, int nbins) {
    if (!op || !op->ctx || !gains || nbins < 1 || nbins > 4096)
        return fail(IMP_EINVAL, "imp_toep_set_precond_radial: bad arguments");
    for (int b = 0; b < nbins; ++b)
        if (!(gains[b] > 0.f && std::isfinite(gains[b])))
            return fail(IMP_EINVAL, "imp_toep_set_precond_radial: gain %d is %g", b, gains[b]);
    ToepPlan &P = op->P;
    imp_ctx *ctx = op->ctx;
    IMP_TRY(ensure_dscratch(ctx, (size_t)(nbins + 1) / 2 + 16));
    float *dg = (float *)ctx->dscratch;
    IMP_HIP(hipMemcpyAsync(dg, gains, nbins * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_precond_radial, dim3(ceil_div(P.nvox, 256)), dim3(256), 0, ctx->stream,
                       P.pw_chan, dg, nbins, P.N1, P.N2, P.N3, P.pw);
    IMP_CHECK_LAUNCH();
    IMP_HIP(hipStreamSynchronize(ctx->stream));
    return IMP_OK;
}

extern "C" int imp_toep_get_schedule(imp_toep *op, int *coils_per_group) {
    if (!op || !coils_per_group) return fail(IMP_EINVAL, "imp_toep_get_schedule: bad arguments");
    *coils_per_group = op->P.coils_per_group;
    return IMP_OK;
}

extern "C" int imp_toep_set_schedule(imp_toep *op, int coils_per_group) {
    if (!op || !op->ctx) return fail(IMP_EINVAL, "imp_toep_set_schedule: bad operator");
    ToepPlan &P = op->P;
    // the partial buffer holds the groups of the schedule chosen at create time
    if (coils_per_group < P.coils_per_group_min || coils_per_group > P.ncoils)
        return fail(IMP_EINVAL, "imp_toep_set_schedule: coils per group %d outside [%d, %d]",
                    coils_per_group, P.coils_per_group_min, P.ncoils);
    P.coils_per_group = coils_per_group;
    return IMP_OK;
}

extern "C" int imp_toep_set_fieldmap(imp_toep *op, const float *fm) {
    if (!op || !op->ctx) return fail(IMP_EINVAL, "imp_toep_set_fieldmap: bad operator");
    ToepPlan &P = op->P;
    hipLaunchKernelGGL(k_fm_to_cycles, dim3(ceil_div(P.nvox, 256)), dim3(256), 0,
                       op->ctx->stream, fm, P.fcyc, P.nvox);
    IMP_CHECK_LAUNCH();
    return IMP_OK;
}

extern "C" int imp_toep_get_precond(imp_toep *op, float *w) {
    if (!op || !op->ctx || !w) return fail(IMP_EINVAL, "imp_toep_get_precond: bad arguments");
    IMP_HIP(hipMemcpyAsync(w, op->P.pw, (size_t)op->P.nvox * sizeof(float),
                           hipMemcpyDeviceToDevice, op->ctx->stream));
    IMP_HIP(hipStreamSynchronize(op->ctx->stream));
    return IMP_OK;
}

extern "C" int imp_toep_precond_apply(imp_toep *op, const float *r_r, const float *r_i,
                                      float *z_r, float *z_i) {
    if (!op || !op->ctx || !r_r || !r_i || !z_r || !z_i)
        return fail(IMP_EINVAL, "imp_toep_precond_apply: bad arguments");
    ToepPlan &P = op->P;
    imp_ctx *ctx = op->ctx;
    IMP_TRY(split_to_c(ctx, r_r, r_i, P.vx, P.nvox));
    IMP_TRY(precond_apply(ctx, P, P.vx, P.vq));
    IMP_TRY(c_to_split(ctx, P.vq, z_r, z_i, P.nvox));
    IMP_HIP(hipStreamSynchronize(ctx->stream));
    return IMP_OK;
}

extern "C" int imp_toep_set_weight(imp_toep *op, const float *w) {
    if (!op) return fail(IMP_EINVAL, "imp_toep_set_weight: NULL operator");
    op->P.w = const_cast<float *>(w);
    return IMP_OK;
}

extern "C" int imp_toep_set_sparse_reg(imp_toep *op, const imp_sparse_reg *reg) {
    if (!op) return fail(IMP_EINVAL, "imp_toep_set_sparse_reg: NULL operator");
    const SparseReg *R = sparse_reg_of(reg);
    if (reg && !R) return fail(IMP_EINVAL, "imp_toep_set_sparse_reg: regularizer has no context");
    // the regularizer must live on the operator's own context: its CSR
    // arrays and work vector are used on that context's stream, and only
    // that context detaches the operator when it is destroyed
    if (R && (!op->ctx || sparse_reg_ctx(reg) != op->ctx))
        return fail(IMP_EINVAL, "imp_toep_set_sparse_reg: regularizer and operator belong to "
                                "different contexts");
    if (R) {
        const long ndim = op->P.N3 > 1 ? 3 : 2;
        if ((long)R->cols != op->P.nvox || (long)R->rows != ndim * op->P.nvox)
            return fail(IMP_EINVAL, "imp_toep_set_sparse_reg: C is %d x %d, expected %ld x %ld",
                        R->rows, R->cols, ndim * op->P.nvox, op->P.nvox);
    }
    op->P.sreg = R;
    op->P.sreg_lost = false;
    return IMP_OK;
}

extern "C" int imp_toep_apply(imp_toep *op, const float *p_r, const float *p_i, float *out_r,
                              float *out_i) {
    if (!op || !op->ctx || !p_r || !p_i || !out_r || !out_i)
        return fail(IMP_EINVAL, "imp_toep_apply: bad arguments");
    ToepPlan &P = op->P;
    imp_ctx *ctx = op->ctx;
    IMP_TRY(split_to_c(ctx, p_r, p_i, P.vx, P.nvox));
    IMP_TRY(toep_apply_full(ctx, P, P.vx, P.vq));
    IMP_TRY(c_to_split(ctx, P.vq, out_r, out_i, P.nvox));
    return IMP_OK;
}

extern "C" int imp_toep_cg(imp_toep *op, const float *b_r, const float *b_i, float *x_r,
                           float *x_i, int max_iter, double tol, int flags,
                           imp_cg_stats *stats) {
    if (!op || !op->ctx || !b_r || !b_i || !x_r || !x_i || max_iter < 0)
        return fail(IMP_EINVAL, "imp_toep_cg: bad arguments");
    ToepPlan &P = op->P;
    imp_ctx *ctx = op->ctx;
    IMP_TRY(split_to_c(ctx, b_r, b_i, P.vb, P.nvox));
    if (flags & IMP_CG_USE_X0) IMP_TRY(split_to_c(ctx, x_r, x_i, P.vx, P.nvox));
    IMP_TRY(toep_cg(ctx, P, P.vb, P.vx, max_iter, tol, flags, stats));
    IMP_TRY(c_to_split(ctx, P.vx, x_r, x_i, P.nvox));
    IMP_HIP(hipStreamSynchronize(ctx->stream));
    return IMP_OK;
}
