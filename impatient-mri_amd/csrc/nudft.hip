// Direct, field-corrected NUDFT on gfx950: forward F, adjoint F^H (with
// SENSE) and the Toeplitz kernels Q_l.
//
// All three are one computation,
//     out[a][c] = sum_b w[b][c] * exp(i * sign * 2pi * (P_a . B_b)),
// with 4-vectors P_a (owned by a lane) and B_b (streamed), because the phase
// 2pi(k.x) + w(x) t is the dot product of (x, y, z, w/2pi) with (kx, ky, kz, t):
//   adjoint  a = voxel  (x,y,z,fm/2pi), b = sample (kx,ky,kz,t), w = conj(phi) d_c
//   forward  a = sample (kx,ky,kz,t),   b = voxel  (x,y,z,fm/2pi), w = S_c rho
//   Q_l      a = Q-grid point (r,0),    b = sample (kx,ky,kz,0),   w = a_l(t) |phi|^2
// The reference evaluates one coil (or one segment l) per pass and recomputes
// the sincos every time (computeFH.cmem.cu:1166-1178, computeQ.cmem.cu:951-963).
// Here one sincos feeds G right-hand sides held in registers, so for 32 coils
// the complex multiply-accumulates dominate instead of the transcendentals.
//
// Streamed operands are wave-uniform, so they are read with scalar loads
// through the constant cache (the role CUDA __constant__ tiles play in the
// reference, ftGpu.cu:103, computeFH.cmem.cu:108-109) without tile relaunches.
//
// Precision: the phase is formed in revolutions with split-precision
// products (two-product via FMA, exact fractional reduction of each product)
// and reduced to [-1/2, 1/2] before the hardware sin/cos, which take
// revolutions.  Sample sums accumulate in FP32 over blocks of BLK samples and
// are folded into a compensated (two-sum) total; chunk partials are added in
// FP64.
#include "common.h"

#include <cmath>
#include <cstdio>
#include <vector>

namespace imp {

namespace {

constexpr int NU_THREADS = 256;
constexpr int NU_BLK = 32;  // samples per FP32 block before compensated fold

enum { EPI_COMBINE = 0, EPI_PERCOIL = 1 };

// Stream stride in floats for G right-hand sides; multiple of 4 so every
// element starts on a 16-byte boundary (one s_load_dwordx4 for B).
__host__ __device__ constexpr int stream_stride(int G, bool realw) {
    return realw ? 4 + ((G + 3) & ~3) : 4 + 2 * ((G + 1) & ~1);
}

template <int G, bool D3, bool REALW, int EPI>
__global__ __launch_bounds__(NU_THREADS) void k_nudft(
    const float4 *__restrict__ lanes, int n_lane,
    const float *__restrict__ stream, int n_stream, int chunk_len,
    float sign, int ncoils,
    const float *__restrict__ s_r, const float *__restrict__ s_i,
    float2 *__restrict__ partial) {
    constexpr int S = stream_stride(G, REALW);
    const int a = blockIdx.x * NU_THREADS + threadIdx.x;
    const int chunk = blockIdx.y;
    const int group = blockIdx.z;
    const int nchunks = gridDim.y;
    const int b0 = chunk * chunk_len;
    const int b1 = min(n_stream, b0 + chunk_len);
    const float *__restrict__ gstream = stream + (size_t)group * n_stream * S;

    const float4 P = lanes[min(a, n_lane - 1)];

    float hr[G], hi[G], lr[G], li[G];
#pragma unroll
    for (int g = 0; g < G; ++g) { hr[g] = hi[g] = lr[g] = li[g] = 0.f; }

    for (int b = b0; b < b1; b += NU_BLK) {
        const int be = min(b + NU_BLK, b1);
        float ar[G], ai[G];
#pragma unroll
        for (int g = 0; g < G; ++g) { ar[g] = ai[g] = 0.f; }
        for (int j = b; j < be; ++j) {
            const float *__restrict__ e = gstream + (size_t)j * S;
            const float4 B = *reinterpret_cast<const float4 *>(e);
            // phase in revolutions, split precision
            const float p1 = P.x * B.x;
            const float e1 = __builtin_fmaf(P.x, B.x, -p1);
            const float p2 = P.y * B.y;
            const float e2 = __builtin_fmaf(P.y, B.y, -p2);
            float h = frac_near(p1) + frac_near(p2);
            float lo = e1 + e2;
            if constexpr (D3) {
                const float p3 = P.z * B.z;
                const float e3 = __builtin_fmaf(P.z, B.z, -p3);
                h += frac_near(p3);
                lo += e3;
            }
            const float u = frac_near(h + __builtin_fmaf(P.w, B.w, lo));
            float sn, cs;
            sincos_rev(u, &sn, &cs);
            sn *= sign;
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if constexpr (REALW) {
                    const float w = e[4 + g];
                    ar[g] = __builtin_fmaf(w, cs, ar[g]);
                    ai[g] = __builtin_fmaf(w, sn, ai[g]);
                } else {
                    const float wr = e[4 + 2 * g];
                    const float wi = e[5 + 2 * g];
                    ar[g] = __builtin_fmaf(wr, cs, __builtin_fmaf(-wi, sn, ar[g]));
                    ai[g] = __builtin_fmaf(wr, sn, __builtin_fmaf(wi, cs, ai[g]));
                }
            }
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            float s, err;
            two_sum(hr[g], ar[g], s, err);
            hr[g] = s;
            lr[g] += err;
            two_sum(hi[g], ai[g], s, err);
            hi[g] = s;
            li[g] += err;
        }
    }
    if (a >= n_lane) return;

    if constexpr (EPI == EPI_COMBINE) {
        // sum_g conj(S_g(a)) * acc_g
        float orr = 0.f, oi = 0.f;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int c = group * G + g;
            if (c < ncoils) {
                const float vr = hr[g] + lr[g];
                const float vi = hi[g] + li[g];
                float sr = 1.f, si = 0.f;
                if (s_r) {
                    sr = s_r[(size_t)c * n_lane + a];
                    si = s_i[(size_t)c * n_lane + a];
                }
                orr += sr * vr + si * vi;
                oi += sr * vi - si * vr;
            }
        }
        partial[((size_t)group * nchunks + chunk) * n_lane + a] = make_float2(orr, oi);
    } else {
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int c = group * G + g;
            if (c < ncoils)
                partial[((size_t)chunk * ncoils + c) * n_lane + a] =
                    make_float2(hr[g] + lr[g], hi[g] + li[g]);
        }
    }
}

// ---- stream / lane packing -------------------------------------------------

// adjoint stream: {kx,ky,kz,t, conj(phi_c) d_c ...} per sample, per coil group
__global__ void k_pack_adjoint(int num_k, int ncoils, int G, int S,
                               const float *kx, const float *ky, const float *kz,
                               const float *t, const float *phi_r,
                               const float *phi_i, const float *d_r,
                               const float *d_i, float *stream) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int group = blockIdx.y;
    if (k >= num_k) return;
    float *e = stream + ((size_t)group * num_k + k) * S;
    e[0] = kx[k];
    e[1] = ky[k];
    e[2] = kz ? kz[k] : 0.f;
    e[3] = t[k];
    for (int j = 4 + 2 * G; j < S; ++j) e[j] = 0.f;
    for (int g = 0; g < G; ++g) {
        const int c = group * G + g;
        float wr = 0.f, wi = 0.f;
        if (c < ncoils) {
            const size_t i = (size_t)c * num_k + k;
            const float dr = d_r[i], di = d_i[i];
            if (phi_r) {
                // conj(phi) * d (ComputeRhoPhiGPU, computeFH.cmem.cu:112-127)
                const float pr = phi_r[i], pi = phi_i[i];
                wr = pr * dr + pi * di;
                wi = pr * di - pi * dr;
            } else {
                wr = dr;
                wi = di;
            }
        }
        e[4 + 2 * g] = wr;
        e[5 + 2 * g] = wi;
    }
}

// forward stream: {x,y,z,fm/2pi, S_c rho ...} per voxel, per coil group
__global__ void k_pack_forward(int num_x, int ncoils, int G, int S,
                               const float *x, const float *y, const float *z,
                               const float *fm, const float *rho_r,
                               const float *rho_i, const float *s_r,
                               const float *s_i, float *stream) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const int group = blockIdx.y;
    if (v >= num_x) return;
    float *e = stream + ((size_t)group * num_x + v) * S;
    e[0] = x[v];
    e[1] = y[v];
    e[2] = z ? z[v] : 0.f;
    e[3] = fm ? (float)((double)fm[v] * (1.0 / (2.0 * M_PI))) : 0.f;
    for (int j = 4 + 2 * G; j < S; ++j) e[j] = 0.f;
    const float rr = rho_r[v], ri = rho_i[v];
    for (int g = 0; g < G; ++g) {
        const int c = group * G + g;
        float wr = 0.f, wi = 0.f;
        if (c < ncoils) {
            float sr = 1.f, si = 0.f;
            if (s_r) {
                sr = s_r[(size_t)c * num_x + v];
                si = s_i[(size_t)c * num_x + v];
            }
            wr = sr * rr - si * ri;
            wi = sr * ri + si * rr;
        }
        e[4 + 2 * g] = wr;
        e[5 + 2 * g] = wi;
    }
}

// Q stream: {kx,ky,kz,0, a_l(t_k)|phi_k|^2 ...} per sample, per segment group.
// a_l is the Hanning interpolator of hanning_d (computeQ.cmem.cu:125-138),
// evaluated in FP64.
__global__ void k_pack_q(int num_k, int nseg, int G, int S, const float *kx,
                         const float *ky, const float *kz, const float *t,
                         const float *phi_r, const float *phi_i, float tau,
                         float t0, float *stream) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int group = blockIdx.y;
    if (k >= num_k) return;
    float *e = stream + ((size_t)group * num_k + k) * S;
    e[0] = kx[k];
    e[1] = ky[k];
    e[2] = kz ? kz[k] : 0.f;
    e[3] = 0.f;
    for (int j = 4 + G; j < S; ++j) e[j] = 0.f;
    double mag = 1.0;
    if (phi_r) {
        const double pr = phi_r[k], pi = phi_i[k];
        mag = pr * pr + pi * pi;
    }
    const double tk = t[k];
    for (int g = 0; g < G; ++g) {
        const int l = group * G + g;
        double w = 0.0;
        if (l < nseg) {
            const double arg = tk - (double)tau * l - (double)t0;
            if (fabs(arg) < (double)tau) w = 0.5 + 0.5 * cos(M_PI * arg / (double)tau);
        }
        e[4 + g] = (float)(w * mag);
    }
}

__global__ void k_lanes_from_arrays(int n, const float *x, const float *y,
                                    const float *z, const float *w, double wscale,
                                    float4 *lanes) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    lanes[i] = make_float4(x[i], y[i], z ? z[i] : 0.f,
                           w ? (float)((double)w[i] * wscale) : 0.f);
}

// standard Q grid (main_mri.cu:461-500): y outer, x, z inner, r = (x/Nx, y/Ny, z/Nz)
__global__ void k_lanes_qgrid(int Nx, int Ny, int Nz, float4 *lanes) {
    const int M2 = 2 * Nx, M3 = Nz > 1 ? 2 * Nz : 1;
    const int n = 2 * Ny * M2 * M3;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int iz = i % M3;
    const int ix = (i / M3) % M2;
    const int iy = i / (M3 * M2);
    const float gx = (float)(ix - Nx) / (float)Nx;
    const float gy = (float)(iy - Ny) / (float)Ny;
    const float gz = Nz > 1 ? (float)(iz - Nz) / (float)Nz : 0.f;
    lanes[i] = make_float4(gx, gy, gz, 0.f);
}

// ---- reductions over chunk partials (FP64) ---------------------------------

__global__ void k_reduce_combine(const float2 *partial, int nparts, int n,
                                 float *out_r, float *out_i) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    double sr = 0.0, si = 0.0;
    for (int p = 0; p < nparts; ++p) {
        const float2 v = partial[(size_t)p * n + a];
        sr += v.x;
        si += v.y;
    }
    out_r[a] = (float)sr;
    out_i[a] = (float)si;
}

// out[c][a] = sum_chunks partial[chunk][c][a]; cumulative: prefix sum over c
__global__ void k_reduce_percoil(const float2 *partial, int nchunks, int ncoils,
                                 int n, int cumulative, float *out_r,
                                 float *out_i) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    double cr = 0.0, ci = 0.0;
    for (int c = 0; c < ncoils; ++c) {
        double sr = 0.0, si = 0.0;
        for (int p = 0; p < nchunks; ++p) {
            const float2 v = partial[((size_t)p * ncoils + c) * n + a];
            sr += v.x;
            si += v.y;
        }
        if (cumulative) {
            cr += sr;
            ci += si;
            sr = cr;
            si = ci;
        }
        out_r[(size_t)c * n + a] = (float)sr;
        out_i[(size_t)c * n + a] = (float)si;
    }
}

// ---- dispatch ----------------------------------------------------------------

using KFn = void (*)(const float4 *, int, const float *, int, int, float, int,
                     const float *, const float *, float2 *);

template <int G, bool REALW, int EPI>
KFn pick_d3(bool d3) {
    return d3 ? (KFn)k_nudft<G, true, REALW, EPI> : (KFn)k_nudft<G, false, REALW, EPI>;
}

template <bool REALW, int EPI>
KFn pick(int G, bool d3) {
    switch (G) {
        case 1: return pick_d3<1, REALW, EPI>(d3);
        case 2: return pick_d3<2, REALW, EPI>(d3);
        case 4: return pick_d3<4, REALW, EPI>(d3);
        case 8: return pick_d3<8, REALW, EPI>(d3);
        case 12: return pick_d3<12, REALW, EPI>(d3);
        default: return pick_d3<16, REALW, EPI>(d3);
    }
}

int group_size(int nrhs) {
    if (nrhs <= 1) return 1;
    if (nrhs <= 2) return 2;
    if (nrhs <= 4) return 4;
    if (nrhs <= 8) return 8;
    if (nrhs <= 12) return 12;
    return 16;
}

// Number of stream chunks so the launch holds enough waves to fill the GPU.
int pick_chunks(imp_ctx *ctx, int lane_blocks, int groups, int n_stream) {
    const long target = 8L * ctx->num_cus;  // workgroups (4 waves each)
    long have = (long)lane_blocks * groups;
    int chunks = 1;
    while (have * chunks < target && n_stream / (chunks * 2) >= 4 * NU_BLK) chunks *= 2;
    return chunks;
}

struct NudftLaunch {
    const float4 *lanes;
    int n_lane;
    const float *stream;
    int n_stream;
    int G, groups, nrhs;
    bool d3, realw;
    int epi;
    float sign;
    const float *s_r, *s_i;
};

int run_nudft(imp_ctx *ctx, const NudftLaunch &L, float2 **partial_out, int *nchunks_out) {
    const int lane_blocks = ceil_div(L.n_lane, NU_THREADS);
    const int chunks = pick_chunks(ctx, lane_blocks, L.groups, L.n_stream);
    int chunk_len = ceil_div(L.n_stream, chunks);
    chunk_len = ceil_div(chunk_len, NU_BLK) * NU_BLK;
    const size_t nparts = L.epi == EPI_COMBINE ? (size_t)L.groups * chunks
                                               : (size_t)chunks * L.nrhs;
    void *part = nullptr;
    IMP_TRY(workspace(ctx, 2, nparts * L.n_lane * sizeof(float2), &part));
    KFn fn = L.realw ? (L.epi == EPI_COMBINE ? pick<true, EPI_COMBINE>(L.G, L.d3)
                                             : pick<true, EPI_PERCOIL>(L.G, L.d3))
                     : (L.epi == EPI_COMBINE ? pick<false, EPI_COMBINE>(L.G, L.d3)
                                             : pick<false, EPI_PERCOIL>(L.G, L.d3));
    dim3 grid(lane_blocks, chunks, L.groups);
    hipLaunchKernelGGL(fn, grid, dim3(NU_THREADS), 0, ctx->stream, L.lanes, L.n_lane,
                       L.stream, L.n_stream, chunk_len, L.sign, L.nrhs, L.s_r, L.s_i,
                       (float2 *)part);
    IMP_CHECK_LAUNCH();
    *partial_out = (float2 *)part;
    *nchunks_out = chunks;
    return IMP_OK;
}

}  // namespace

// ---- operator entry points (device pointers) --------------------------------

int nudft_adjoint(imp_ctx *ctx, int num_k, int num_x, int ncoils, const float *kx,
                  const float *ky, const float *kz, const float *t, const float *x,
                  const float *y, const float *z, const float *fm,
                  const float *phi_r, const float *phi_i, const float *d_r,
                  const float *d_i, const float *s_r, const float *s_i,
                  float *out_r, float *out_i) {
    if (!ctx || num_k <= 0 || num_x <= 0 || ncoils <= 0 || !kx || !ky || !t ||
        !x || !y || !d_r || !d_i || !out_r || !out_i || (s_r && !s_i) ||
        (phi_r && !phi_i))
        return fail(IMP_EINVAL, "nudft_adjoint: bad arguments");
    const bool d3 = kz && z;
    const int G = group_size(ncoils);
    const int groups = ceil_div(ncoils, G);
    const int S = stream_stride(G, false);
    void *stream = nullptr, *lanes = nullptr;
    IMP_TRY(workspace(ctx, 0, (size_t)groups * num_k * S * sizeof(float), &stream));
    IMP_TRY(workspace(ctx, 1, (size_t)num_x * sizeof(float4), &lanes));
    hipLaunchKernelGGL(k_pack_adjoint, dim3(ceil_div(num_k, 256), groups), dim3(256), 0,
                       ctx->stream, num_k, ncoils, G, S, kx, ky, d3 ? kz : nullptr, t,
                       phi_r, phi_i, d_r, d_i, (float *)stream);
    IMP_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_lanes_from_arrays, dim3(ceil_div(num_x, 256)), dim3(256), 0,
                       ctx->stream, num_x, x, y, d3 ? z : nullptr, fm,
                       1.0 / (2.0 * M_PI), (float4 *)lanes);
    IMP_CHECK_LAUNCH();
    NudftLaunch L{(const float4 *)lanes, num_x, (const float *)stream, num_k, G, groups,
                  ncoils, d3, false, EPI_COMBINE, +1.f, s_r, s_i};
    float2 *partial = nullptr;
    int chunks = 0;
    IMP_TRY(run_nudft(ctx, L, &partial, &chunks));
    hipLaunchKernelGGL(k_reduce_combine, dim3(ceil_div(num_x, 256)), dim3(256), 0,
                       ctx->stream, partial, groups * chunks, num_x, out_r, out_i);
    IMP_CHECK_LAUNCH();
    return IMP_OK;
}

int nudft_forward(imp_ctx *ctx, int num_k, int num_x, int ncoils, const float *kx,
                  const float *ky, const float *kz, const float *t, const float *x,
                  const float *y, const float *z, const float *fm,
                  const float *rho_r, const float *rho_i, const float *s_r,
                  const float *s_i, float *d_r, float *d_i) {
    if (!ctx || num_k <= 0 || num_x <= 0 || ncoils <= 0 || !kx || !ky || !t ||
        !x || !y || !rho_r || !rho_i || !d_r || !d_i || (s_r && !s_i))
        return fail(IMP_EINVAL, "nudft_forward: bad arguments");
    const bool d3 = kz && z;
    const int G = group_size(ncoils);
    const int groups = ceil_div(ncoils, G);
    const int S = stream_stride(G, false);
    void *stream = nullptr, *lanes = nullptr;
    IMP_TRY(workspace(ctx, 0, (size_t)groups * num_x * S * sizeof(float), &stream));
    IMP_TRY(workspace(ctx, 1, (size_t)num_k * sizeof(float4), &lanes));
    hipLaunchKernelGGL(k_pack_forward, dim3(ceil_div(num_x, 256), groups), dim3(256), 0,
                       ctx->stream, num_x, ncoils, G, S, x, y, d3 ? z : nullptr, fm,
                       rho_r, rho_i, s_r, s_i, (float *)stream);
    IMP_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_lanes_from_arrays, dim3(ceil_div(num_k, 256)), dim3(256), 0,
                       ctx->stream, num_k, kx, ky, d3 ? kz : nullptr, t, 1.0,
                       (float4 *)lanes);
    IMP_CHECK_LAUNCH();
    NudftLaunch L{(const float4 *)lanes, num_k, (const float *)stream, num_x, G, groups,
                  ncoils, d3, false, EPI_PERCOIL, -1.f, nullptr, nullptr};
    float2 *partial = nullptr;
    int chunks = 0;
    IMP_TRY(run_nudft(ctx, L, &partial, &chunks));
    hipLaunchKernelGGL(k_reduce_percoil, dim3(ceil_div(num_k, 256)), dim3(256), 0,
                       ctx->stream, partial, chunks, ncoils, num_k, 0, d_r, d_i);
    IMP_CHECK_LAUNCH();
    return IMP_OK;
}

int toeplitz_q(imp_ctx *ctx, int num_k, const float *kx, const float *ky,
               const float *kz, const float *t, const float *phi_r,
               const float *phi_i, int Nx, int Ny, int Nz, int num_q,
               const float *gx, const float *gy, const float *gz, int ntime_segs,
               int flags, const float *t_host_first_last, float *q_r, float *q_i) {
    if (!ctx || num_k <= 0 || !kx || !ky || !t || ntime_segs < 0 || !q_r || !q_i ||
        (phi_r && !phi_i) || ((gx == nullptr) != (gy == nullptr)))
        return fail(IMP_EINVAL, "toeplitz_q: bad arguments");
    const bool std_grid = gx == nullptr;
    if (std_grid) {
        if (Nx <= 0 || Ny <= 0 || Nz <= 0)
            return fail(IMP_EINVAL, "toeplitz_q: bad grid size");
        const long M3 = Nz > 1 ? 2L * Nz : 1L;
        const long nq = 4L * Nx * Ny * M3;
        if (nq > 0x7fffffffL) return fail(IMP_EINVAL, "toeplitz_q: grid too large");
        num_q = (int)nq;
    } else if (num_q <= 0) {
        return fail(IMP_EINVAL, "toeplitz_q: num_q must be positive");
    }
    const bool d3 = kz && (std_grid ? Nz > 1 : gz != nullptr);
    const int nseg = ntime_segs + 1;
    // tau and t0 as the reference computes them, in FP32 on the host
    // (computeQ.cmem.cu:870-871)
    const float t0 = t_host_first_last[0];
    const float tlast = t_host_first_last[1];
    const float tau = ntime_segs > 0 ? (tlast - t0) / (float)ntime_segs : (tlast - t0);
    const int G = group_size(nseg);
    const int groups = ceil_div(nseg, G);
    const int S = stream_stride(G, true);
    void *stream = nullptr, *lanes = nullptr;
    IMP_TRY(workspace(ctx, 0, (size_t)groups * num_k * S * sizeof(float), &stream));
    IMP_TRY(workspace(ctx, 1, (size_t)num_q * sizeof(float4), &lanes));
    hipLaunchKernelGGL(k_pack_q, dim3(ceil_div(num_k, 256), groups), dim3(256), 0,
                       ctx->stream, num_k, nseg, G, S, kx, ky, d3 ? kz : nullptr, t,
                       phi_r, phi_i, tau, t0, (float *)stream);
    IMP_CHECK_LAUNCH();
    if (std_grid) {
        hipLaunchKernelGGL(k_lanes_qgrid, dim3(ceil_div(num_q, 256)), dim3(256), 0,
                           ctx->stream, Nx, Ny, d3 ? Nz : 1, (float4 *)lanes);
    } else {
        hipLaunchKernelGGL(k_lanes_from_arrays, dim3(ceil_div(num_q, 256)), dim3(256), 0,
                           ctx->stream, num_q, gx, gy, d3 ? gz : nullptr, nullptr, 0.0,
                           (float4 *)lanes);
    }
    IMP_CHECK_LAUNCH();
    NudftLaunch L{(const float4 *)lanes, num_q, (const float *)stream, num_k, G, groups,
                  nseg, d3, true, EPI_PERCOIL, +1.f, nullptr, nullptr};
    float2 *partial = nullptr;
    int chunks = 0;
    IMP_TRY(run_nudft(ctx, L, &partial, &chunks));
    hipLaunchKernelGGL(k_reduce_percoil, dim3(ceil_div(num_q, 256)), dim3(256), 0,
                       ctx->stream, partial, chunks, nseg, num_q,
                       (flags & IMP_Q_CUMULATIVE) ? 1 : 0, q_r, q_i);
    IMP_CHECK_LAUNCH();
    return IMP_OK;
}

}  // namespace imp
