// Stockham FFTs in LDS and the Toeplitz normal-operator passes built on them.
//
// The reference applies T_l = F^H diag(a_l) F as
//   crop(IFFT(FFT(ifftshift(pad_center(v))) .* Qf_l)) / G
// with cuFFT on the 2x grid plus separate oversample / fftshift / multiply /
// undersample kernels (recon.cu:1113-1202).  Because circular convolution is
// shift invariant, padding v into the low corner of the 2x grid and cropping
// the low corner gives the same result with no shifts.  The zero half of the
// padded input and the discarded half of the output are pruned per axis:
//   pass A  FFT along the last axis of the N-sized lines (input: S_c e^{-iw tau_l} p)
//   pass B  FFT along x                                  (3D only)
//   pass C  FFT along y, times Qf_l, inverse FFT along y, keep y < N1
//   pass D  inverse FFT along x, keep x < N2             (3D only)
//   pass E  inverse FFT along the last axis, keep < N, times conj(S_c)/G,
//           partial sums over coil groups
//   pass F  out (+)= e^{+iw tau_l} * sum of coil-group partials
// Each pass reads and writes its lines once; Qf_l stays in HBM.
#include "common.h"
#include "toeplitz.h"

#include <cmath>
#include <vector>

namespace imp {

namespace {

constexpr int FT = 256;  // threads per FFT workgroup

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 rot90(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2 *v) {
    if constexpr (R == 2) {
        const float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else {  // R == 4
        const float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        const float2 s13 = cadd(v[1], v[3]), d13 = rot90<INV>(csub(v[1], v[3]));
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        v[1] = cadd(d02, d13);
        v[3] = csub(d02, d13);
    }
}

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

// One Stockham radix-R stage over NL lines of length L held in LDS (line
// stride LS), register staged: read all, barrier, write all, barrier.
// tw[m] = exp(-2 pi i m / L).
template <int L, int NL, int R, bool INV>
__device__ __forceinline__ void stockham_stage(float2 *buf, int LS, int Ns,
                                               const float2 *__restrict__ tw) {
    constexpr int Q = L / R;
    constexpr int NB = NL * Q / FT;
    static_assert(NB >= 1 && NL * Q % FT == 0, "stage too small for the workgroup");
    float2 v[NB][R];
    const int tid = threadIdx.x;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int bf = tid + b * FT;
        const int line = bf / Q, j = bf % Q;
        const float2 *x = buf + line * LS;
#pragma unroll
        for (int r = 0; r < R; ++r) v[b][r] = x[j + r * Q];
        const int k = j & (Ns - 1);
        const int tstep = (L / (Ns * R)) * k;
#pragma unroll
        for (int r = 1; r < R; ++r) {
            float2 w = tw[r * tstep];
            if (INV) w.y = -w.y;
            v[b][r] = cmul(v[b][r], w);
        }
        dft<R, INV>(v[b]);
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int bf = tid + b * FT;
        const int line = bf / Q, j = bf % Q;
        const int k = j & (Ns - 1);
        float2 *y = buf + line * LS + (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) y[r * Ns] = v[b][r];
    }
    __syncthreads();
}

// Full FFT of NL lines in LDS.  Caller has synchronised after filling buf.
template <int L, int NL, bool INV>
__device__ __forceinline__ void fft_lds(float2 *buf, int LS, const float2 *__restrict__ tw) {
    constexpr int LOG = ilog2(L);
    int Ns = 1;
    if constexpr (LOG & 1) {
        stockham_stage<L, NL, 2, INV>(buf, LS, Ns, tw);
        Ns = 2;
    }
    if constexpr (LOG >= 2) {
#pragma unroll
        for (int s = 0; s < LOG / 2; ++s) {
            stockham_stage<L, NL, 4, INV>(buf, LS, Ns, tw);
            Ns *= 4;
        }
    }
}

template <int L>
constexpr int contig_lines() { return (4 * FT) / L > 1 ? (4 * FT) / L : 1; }

template <int L>
constexpr int panel_lines() {
    return L >= 1024 ? 8 : ((4 * FT) / L > 16 ? (4 * FT) / L : 16);
}

// ---------------------------------------------------------------- generic
// In-place FFT of `nlines` contiguous lines of length L.
template <int L, bool INV>
__global__ __launch_bounds__(FT) void k_fft_contig(float2 *data, long nlines,
                                                   const float2 *__restrict__ tw) {
    constexpr int NL = contig_lines<L>();
    __shared__ float2 buf[NL * L];
    const long line0 = (long)blockIdx.x * NL;
    for (int e = threadIdx.x; e < NL * L; e += FT) {
        const long line = line0 + e / L;
        buf[e] = line < nlines ? data[line * L + (e % L)] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft_lds<L, NL, INV>(buf, L, tw);
    for (int e = threadIdx.x; e < NL * L; e += FT) {
        const long line = line0 + e / L;
        if (line < nlines) data[line * L + (e % L)] = buf[e];
    }
}

// Strided lines: line j of batch b starts at b*bstride + j, element i at
// + i*es.  Input elements i >= n_in are zero; outputs i < n_out are stored.
// MODE 0: forward, 1: inverse, 2: forward, multiply by qf (same addressing
// with batch stride 0), inverse.
template <int L, int MODE>
__global__ __launch_bounds__(FT) void k_fft_panel(float2 *data, long bstride, long es,
                                                  int nlines, int n_in, int n_out,
                                                  const float2 *__restrict__ qf,
                                                  const float2 *__restrict__ tw) {
    constexpr int P = panel_lines<L>();
    constexpr int LS = L + 1;
    __shared__ float2 buf[P * LS];
    const int j0 = blockIdx.x * P;
    float2 *base = data + (long)blockIdx.y * bstride;
    for (int e = threadIdx.x; e < P * L; e += FT) {
        const int i = e / P, j = e % P;
        float2 v = make_float2(0.f, 0.f);
        if (i < n_in && j0 + j < nlines) v = base[i * es + j0 + j];
        buf[j * LS + i] = v;
    }
    __syncthreads();
    fft_lds<L, P, MODE == 1>(buf, LS, tw);
    if constexpr (MODE == 2) {
        for (int e = threadIdx.x; e < P * L; e += FT) {
            const int i = e / P, j = e % P;
            if (j0 + j < nlines) buf[j * LS + i] = cmul(buf[j * LS + i], qf[i * es + j0 + j]);
        }
        __syncthreads();
        fft_lds<L, P, true>(buf, LS, tw);
    }
    for (int e = threadIdx.x; e < P * L; e += FT) {
        const int i = e / P, j = e % P;
        if (i < n_out && j0 + j < nlines) base[i * es + j0 + j] = buf[j * LS + i];
    }
}

// ---------------------------------------------------------------- pass A
// Lines (c, o): image line o (N1*N2o lines of nlast elements, contiguous),
// v = S_c * e^{-i 2pi f tau_l} * p, zero-padded to L, forward FFT, written
// to G1 line ((c*N1 + y)*A + x)*L with y = o / N2o, x = o % N2o.
template <int L>
__global__ __launch_bounds__(FT) void k_toep_pass_a(
    const float2 *__restrict__ p, const float2 *__restrict__ sens,
    const float *__restrict__ f, float tau_l, int nimg_lines, int nlast, int N1,
    int N2o, int A, int ncoils, float2 *__restrict__ g1, const float2 *__restrict__ tw) {
    constexpr int NL = contig_lines<L>();
    __shared__ float2 buf[NL * L];
    const long nvox = (long)nimg_lines * nlast;
    const long line0 = (long)blockIdx.x * NL;  // flattened (c, o)
    const long nlines = (long)ncoils * nimg_lines;
    for (int e = threadIdx.x; e < NL * L; e += FT) {
        const long line = line0 + e / L;
        const int i = e % L;
        float2 v = make_float2(0.f, 0.f);
        if (line < nlines && i < nlast) {
            const int c = (int)(line / nimg_lines);
            const long o = line % nimg_lines;
            const long idx = o * nlast + i;
            float sn, cs;
            sincos_rev(frac_near(f[idx] * tau_l), &sn, &cs);
            const float2 pv = p[idx];
            // e^{-i 2pi f tau} p
            float2 u = make_float2(pv.x * cs + pv.y * sn, pv.y * cs - pv.x * sn);
            if (sens) u = cmul(sens[(long)c * nvox + idx], u);
            v = u;
        }
        buf[e] = v;
    }
    __syncthreads();
    fft_lds<L, NL, false>(buf, L, tw);
    for (int e = threadIdx.x; e < NL * L; e += FT) {
        const long line = line0 + e / L;
        if (line >= nlines) continue;
        const int c = (int)(line / nimg_lines);
        const long o = line % nimg_lines;
        const long y = o / N2o, x = o % N2o;
        g1[(((long)c * N1 + y) * A + x) * L + (e % L)] = buf[e];
    }
}

// ---------------------------------------------------------------- pass E
// For NL image lines and the coils of one coil group: inverse FFT of the G1
// lines, keep the first nlast elements, sum conj(S_c) * z_c over the group,
// write partial[group][voxel].
template <int L>
__global__ __launch_bounds__(FT) void k_toep_pass_e(
    const float2 *__restrict__ g1, const float2 *__restrict__ sens, int nimg_lines,
    int nlast, int N1, int N2o, int A, int ncoils, int coils_per_group,
    float2 *__restrict__ partial, const float2 *__restrict__ tw) {
    constexpr int NL = contig_lines<L>();
    constexpr int PER = (NL * L + FT - 1) / FT;
    __shared__ float2 buf[NL * L];
    const long nvox = (long)nimg_lines * nlast;
    const long o0 = (long)blockIdx.x * NL;
    const int group = blockIdx.y;
    const int c0 = group * coils_per_group;
    const int c1 = min(ncoils, c0 + coils_per_group);
    float2 acc[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) acc[q] = make_float2(0.f, 0.f);
    for (int c = c0; c < c1; ++c) {
        for (int e = threadIdx.x; e < NL * L; e += FT) {
            const long o = o0 + e / L;
            float2 v = make_float2(0.f, 0.f);
            if (o < nimg_lines) {
                const long y = o / N2o, x = o % N2o;
                v = g1[(((long)c * N1 + y) * A + x) * L + (e % L)];
            }
            buf[e] = v;
        }
        __syncthreads();
        fft_lds<L, NL, true>(buf, L, tw);
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = threadIdx.x + q * FT;
            if (e < NL * L) {
                const long o = o0 + e / L;
                const int i = e % L;
                if (o < nimg_lines && i < nlast) {
                    const float2 z = buf[e];
                    if (sens) {
                        const float2 s = sens[(long)c * nvox + o * nlast + i];
                        // conj(s) * z
                        acc[q].x += s.x * z.x + s.y * z.y;
                        acc[q].y += s.x * z.y - s.y * z.x;
                    } else {
                        acc[q].x += z.x;
                        acc[q].y += z.y;
                    }
                }
            }
        }
        __syncthreads();  // buf reused by the next coil
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = threadIdx.x + q * FT;
        if (e < NL * L) {
            const long o = o0 + e / L;
            const int i = e % L;
            if (o < nimg_lines && i < nlast)
                partial[(long)group * nvox + o * nlast + i] = acc[q];
        }
    }
}

// ---------------------------------------------------------------- pass F
// out (+)= scale * e^{+i 2pi f tau_l} * sum_groups partial[g]
__global__ void k_toep_pass_f(const float2 *__restrict__ partial, int ngroups, long nvox,
                              const float *__restrict__ f, float tau_l, float scale,
                              int accumulate, float2 *__restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nvox) return;
    float2 s = make_float2(0.f, 0.f);
    for (int g = 0; g < ngroups; ++g) {
        const float2 v = partial[(long)g * nvox + i];
        s.x += v.x;
        s.y += v.y;
    }
    float sn, cs;
    sincos_rev(frac_near(f[i] * tau_l), &sn, &cs);
    float2 r = make_float2(scale * (s.x * cs - s.y * sn), scale * (s.x * sn + s.y * cs));
    if (accumulate) {
        const float2 o = out[i];
        r.x += o.x;
        r.y += o.y;
    }
    out[i] = r;
}

// ifftshift of Q_l (split, standard grid) into interleaved Qf: out[m] = Q[(m + M/2) mod M]
__global__ void k_ifftshift_q(const float *__restrict__ qr, const float *__restrict__ qi,
                              int M1, int M2, int M3, float2 *__restrict__ out) {
    const long n = (long)M1 * M2 * M3;
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const int m3 = (int)(m % M3);
    const int m2 = (int)((m / M3) % M2);
    const int m1 = (int)(m / ((long)M3 * M2));
    const int s1 = (m1 + M1 / 2) % M1;
    const int s2 = (m2 + M2 / 2) % M2;
    const int s3 = M3 > 1 ? (m3 + M3 / 2) % M3 : 0;
    const long src = ((long)s1 * M2 + s2) * M3 + s3;
    out[m] = make_float2(qr[src], qi[src]);
}

// ---------------------------------------------------------------- dispatch
#define IMP_L_SWITCH(Lv, BODY)                 \
    switch (Lv) {                              \
        case 2: { constexpr int L_ = 2; BODY; } break;       \
        case 4: { constexpr int L_ = 4; BODY; } break;       \
        case 8: { constexpr int L_ = 8; BODY; } break;       \
        case 16: { constexpr int L_ = 16; BODY; } break;     \
        case 32: { constexpr int L_ = 32; BODY; } break;     \
        case 64: { constexpr int L_ = 64; BODY; } break;     \
        case 128: { constexpr int L_ = 128; BODY; } break;   \
        case 256: { constexpr int L_ = 256; BODY; } break;   \
        case 512: { constexpr int L_ = 512; BODY; } break;   \
        case 1024: { constexpr int L_ = 1024; BODY; } break; \
        case 2048: { constexpr int L_ = 2048; BODY; } break; \
        default: return fail(IMP_EUNSUPPORTED, "FFT length %d unsupported", Lv); \
    }

bool fft_len_ok(int L) { return L >= 2 && L <= 2048 && (L & (L - 1)) == 0; }

}  // namespace

// ------------------------------------------------------------ twiddles
int twiddles(imp_ctx *ctx, int L, const float2 **out) {
    // table for length L lives at offset L of a 4096-entry buffer
    if (!fft_len_ok(L)) return fail(IMP_EUNSUPPORTED, "FFT length %d unsupported", L);
    if (!ctx->twiddle) {
        std::vector<float2> h(4096, make_float2(0.f, 0.f));
        for (int n = 2; n <= 2048; n *= 2)
            for (int m = 0; m < n; ++m) {
                const double a = -2.0 * M_PI * (double)m / (double)n;
                h[n + m] = make_float2((float)cos(a), (float)sin(a));
            }
        IMP_HIP(hipMalloc(&ctx->twiddle, 4096 * sizeof(float2)));
        IMP_HIP(hipMemcpy(ctx->twiddle, h.data(), 4096 * sizeof(float2), hipMemcpyHostToDevice));
    }
    *out = ctx->twiddle + L;
    return IMP_OK;
}

// In-place full FFT of an interleaved [M1][M2][M3] grid (M3 = 1 for 2D).
int fft_grid(imp_ctx *ctx, float2 *data, int M1, int M2, int M3, bool inverse) {
    const float2 *tw = nullptr;
    const bool d3 = M3 > 1;
    const int Ml = d3 ? M3 : M2;
    // last axis: contiguous lines
    IMP_TRY(twiddles(ctx, Ml, &tw));
    const long nl = (long)M1 * (d3 ? M2 : 1);
    IMP_L_SWITCH(Ml, {
        const int NL = contig_lines<L_>();
        if (inverse)
            hipLaunchKernelGGL((k_fft_contig<L_, true>), dim3(ceil_div(nl, NL)), dim3(FT), 0,
                               ctx->stream, data, nl, tw);
        else
            hipLaunchKernelGGL((k_fft_contig<L_, false>), dim3(ceil_div(nl, NL)), dim3(FT), 0,
                               ctx->stream, data, nl, tw);
    });
    IMP_CHECK_LAUNCH();
    if (d3) {  // x axis: batch y, lines z
        IMP_TRY(twiddles(ctx, M2, &tw));
        IMP_L_SWITCH(M2, {
            const int P = panel_lines<L_>();
            dim3 grid(ceil_div(M3, P), M1);
            if (inverse)
                hipLaunchKernelGGL((k_fft_panel<L_, 1>), grid, dim3(FT), 0, ctx->stream, data,
                                   (long)M2 * M3, (long)M3, M3, M2, M2, nullptr, tw);
            else
                hipLaunchKernelGGL((k_fft_panel<L_, 0>), grid, dim3(FT), 0, ctx->stream, data,
                                   (long)M2 * M3, (long)M3, M3, M2, M2, nullptr, tw);
        });
        IMP_CHECK_LAUNCH();
    }
    // y axis: one batch, lines = M2*M3
    IMP_TRY(twiddles(ctx, M1, &tw));
    const int nlines = M2 * M3;
    IMP_L_SWITCH(M1, {
        const int P = panel_lines<L_>();
        dim3 grid(ceil_div(nlines, P), 1);
        if (inverse)
            hipLaunchKernelGGL((k_fft_panel<L_, 1>), grid, dim3(FT), 0, ctx->stream, data, 0L,
                               (long)nlines, nlines, M1, M1, nullptr, tw);
        else
            hipLaunchKernelGGL((k_fft_panel<L_, 0>), grid, dim3(FT), 0, ctx->stream, data, 0L,
                               (long)nlines, nlines, M1, M1, nullptr, tw);
    });
    IMP_CHECK_LAUNCH();
    return IMP_OK;
}

// Qf_l = FFT(ifftshift(Q_l)) for all l, into plan-owned HBM.
int toep_prepare_qf(imp_ctx *ctx, ToepPlan &P, const float *q_r, const float *q_i) {
    const long G = P.G;
    for (int l = 0; l < P.nseg; ++l) {
        float2 *dst = P.qf + (long)l * G;
        hipLaunchKernelGGL(k_ifftshift_q, dim3(ceil_div(G, 256)), dim3(256), 0, ctx->stream,
                           q_r + (long)l * G, q_i + (long)l * G, P.M1, P.M2, P.M3, dst);
        IMP_CHECK_LAUNCH();
        IMP_TRY(fft_grid(ctx, dst, P.M1, P.M2, P.M3, false));
    }
    return IMP_OK;
}

// out = sum_c conj(S_c) sum_l e^{+iw tau_l} T_l (e^{-iw tau_l} S_c p)   (interleaved, device)
int toep_apply_normal(imp_ctx *ctx, ToepPlan &P, const float2 *p, float2 *out) {
    const bool d3 = P.N3 > 1;
    const int N2o = d3 ? P.N2 : 1;
    const int A = d3 ? P.M2 : 1;
    const int nlast = d3 ? P.N3 : P.N2;
    const int Ml = d3 ? P.M3 : P.M2;
    const int nimg_lines = P.N1 * N2o;
    const long nvox = P.nvox;
    const float scale = (float)(1.0 / (double)P.G);
    const float2 *tw_last = nullptr, *tw_x = nullptr, *tw_y = nullptr;
    IMP_TRY(twiddles(ctx, Ml, &tw_last));
    if (d3) IMP_TRY(twiddles(ctx, P.M2, &tw_x));
    IMP_TRY(twiddles(ctx, P.M1, &tw_y));
    const int C = P.ncoils;
    const int cpg = P.coils_per_group;
    const int ngroups = ceil_div(C, cpg);
    for (int l = 0; l < P.nseg; ++l) {
        const float tau_l = P.t0 + P.tau * (float)l;
        // pass A
        int rec = prof_begin(ctx, "toep_pass_a");
        IMP_L_SWITCH(Ml, {
            const int NL = contig_lines<L_>();
            const long nlines = (long)C * nimg_lines;
            hipLaunchKernelGGL((k_toep_pass_a<L_>), dim3(ceil_div(nlines, NL)), dim3(FT), 0,
                               ctx->stream, p, P.sens, P.fcyc, tau_l, nimg_lines, nlast, P.N1,
                               N2o, A, C, P.g1, tw_last);
        });
        prof_end(ctx, rec);
        IMP_CHECK_LAUNCH();
        // pass B (3D): x axis, batches (c, y), lines z in [0, M3), input x < N2
        if (d3) {
            rec = prof_begin(ctx, "toep_pass_b");
            IMP_L_SWITCH(P.M2, {
                const int Pn = panel_lines<L_>();
                dim3 grid(ceil_div(P.M3, Pn), C * P.N1);
                hipLaunchKernelGGL((k_fft_panel<L_, 0>), grid, dim3(FT), 0, ctx->stream, P.g1,
                                   (long)P.M2 * P.M3, (long)P.M3, P.M3, P.N2, P.M2, nullptr,
                                   tw_x);
            });
            prof_end(ctx, rec);
            IMP_CHECK_LAUNCH();
        }
        // pass C: y axis with Qf multiply, batches c, lines (x,z) in [0, A*Ml)
        {
            const int nl = A * Ml;
            const float2 *qf = P.qf + (long)l * P.G;
            rec = prof_begin(ctx, "toep_pass_c");
            IMP_L_SWITCH(P.M1, {
                const int Pn = panel_lines<L_>();
                dim3 grid(ceil_div(nl, Pn), C);
                hipLaunchKernelGGL((k_fft_panel<L_, 2>), grid, dim3(FT), 0, ctx->stream, P.g1,
                                   (long)P.N1 * nl, (long)nl, nl, P.N1, P.N1, qf, tw_y);
            });
            prof_end(ctx, rec);
            IMP_CHECK_LAUNCH();
        }
        // pass D (3D): inverse x, keep x < N2
        if (d3) {
            rec = prof_begin(ctx, "toep_pass_d");
            IMP_L_SWITCH(P.M2, {
                const int Pn = panel_lines<L_>();
                dim3 grid(ceil_div(P.M3, Pn), C * P.N1);
                hipLaunchKernelGGL((k_fft_panel<L_, 1>), grid, dim3(FT), 0, ctx->stream, P.g1,
                                   (long)P.M2 * P.M3, (long)P.M3, P.M3, P.M2, P.N2, nullptr,
                                   tw_x);
            });
            prof_end(ctx, rec);
            IMP_CHECK_LAUNCH();
        }
        // pass E: inverse last axis, conj(S_c), partial sums over coil groups
        rec = prof_begin(ctx, "toep_pass_e");
        IMP_L_SWITCH(Ml, {
            const int NL = contig_lines<L_>();
            dim3 grid(ceil_div(nimg_lines, NL), ngroups);
            hipLaunchKernelGGL((k_toep_pass_e<L_>), grid, dim3(FT), 0, ctx->stream, P.g1,
                               P.sens, nimg_lines, nlast, P.N1, N2o, A, C, cpg, P.partial,
                               tw_last);
        });
        prof_end(ctx, rec);
        IMP_CHECK_LAUNCH();
        // pass F
        rec = prof_begin(ctx, "toep_pass_f");
        hipLaunchKernelGGL(k_toep_pass_f, dim3(ceil_div(nvox, 256)), dim3(256), 0, ctx->stream,
                           P.partial, ngroups, nvox, P.fcyc, tau_l, scale, l > 0 ? 1 : 0, out);
        prof_end(ctx, rec);
        IMP_CHECK_LAUNCH();
    }
    return IMP_OK;
}

bool toep_sizes_supported(int N1, int N2, int N3) {
    if (!fft_len_ok(2 * N1) || !fft_len_ok(2 * N2)) return false;
    if (N3 > 1 && !fft_len_ok(2 * N3)) return false;
    return true;
}

}  // namespace imp
