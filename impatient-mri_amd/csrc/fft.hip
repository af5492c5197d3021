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
#include "fftcore.h"
#include "toeplitz.h"

#include <cmath>
#include <type_traits>
#include <vector>

namespace imp {

namespace {

using fftc::cmul;

// Stages S .. num_stages - 2, each LDS -> LDS (in place, so a barrier
// between its loads and its stores and one after its stores).
template <int L, int S, bool INV, class LLD, class LST>
__device__ __forceinline__ void middle_stages(int tl, const float2 *__restrict__ tw, LLD lld,
                                              LST lst) {
    if constexpr (S < fftc::num_stages(L) - 1) {
        fftc::stage<L, S, INV, false>(tl, tw, lld, [] { __syncthreads(); }, lst);
        __syncthreads();
        middle_stages<L, S + 1, INV>(tl, tw, lld, lst);
    }
}

// Lines of up to 1024 points read their twiddles from LDS: every stage after
// the first reads one per butterfly input, and an LDS read waits a fraction
// of an L1 hit (1024-point columns with the Qf product: 130 -> 109 us per
// launch, bit-identical).  Every thread of the workgroup calls this.
template <int L>
constexpr bool tw_in_lds() { return fftc::num_stages(L) > 1 && L <= 1024; }
template <int L>
__device__ __forceinline__ const float2 *stage_twiddles(const float2 *tw, float2 *lds) {
    if constexpr (tw_in_lds<L>()) {
        for (int i = threadIdx.x; i < L; i += blockDim.x) lds[i] = tw[i];
        __syncthreads();
        return lds;
    }
    return tw;
}

// Runs all stages of one transform.  FIRST / LAST are the first stage's
// loads and the last stage's stores (global memory or registers); FMID runs
// before the first stage's stores (a barrier when the LDS lines may still be
// read by an earlier transform); LLD / LST address the lines in LDS.
// Every thread of the workgroup must call this (it holds barriers).
template <int L, bool INV, class FIRST, class FMID, class LAST, class LLD, class LST>
__device__ __forceinline__ void run_fft(int tl, const float2 *__restrict__ tw, FIRST first,
                                        FMID fmid, LAST last, LLD lld, LST lst) {
    constexpr int S = fftc::num_stages(L);
    auto nop = [] {};
    if constexpr (S == 1) {
        fftc::stage<L, 0, INV, true>(tl, tw, first, fmid, last);
    } else {
        fftc::stage<L, 0, INV, true>(tl, tw, first, fmid, lst);
        __syncthreads();
        middle_stages<L, 1, INV>(tl, tw, lld, lst);
        fftc::stage<L, S - 1, INV, true>(tl, tw, lld, nop, last);
    }
}

// ---------------------------------------------------------------- rows
// Contiguous lines: T = L/16 consecutive threads per line, so each global
// load or store of a stage covers T consecutive elements.
template <int L>
constexpr int row_lines() { return fftc::row_threads(L) / fftc::threads_per_line(L); }
template <int L>
constexpr int row_lds() { return fftc::num_stages(L) > 1 ? row_lines<L>() * fftc::line_lds(L) : 1; }

// In-place FFT of `nlines` contiguous lines of length L.
template <int L, bool INV>
__global__ __launch_bounds__(fftc::row_threads(L)) void k_fft_rows(float2 *data, long nlines,
                                                         const float2 *__restrict__ tw) {
    constexpr int T = fftc::threads_per_line(L);
    __shared__ float2 buf[row_lds<L>()];
    __shared__ float2 twl[tw_in_lds<L>() ? L : 1];
    tw = stage_twiddles<L>(tw, twl);
    const int tl = threadIdx.x % T, ln = threadIdx.x / T;
    const long line = (long)blockIdx.x * row_lines<L>() + ln;
    const bool ok = line < nlines;
    float2 *src = data + (ok ? line : 0) * L;
    float2 *lb = buf + (fftc::num_stages(L) > 1 ? ln * fftc::line_lds(L) : 0);
    run_fft<L, INV>(
        tl, tw, [&](int, int, int i) { return ok ? src[i] : make_float2(0.f, 0.f); }, [] {},
        [&](int, int, int i, float2 v) {
            if (ok) src[i] = v;
        },
        [&](int, int, int i) { return lb[fftc::lpad(i)]; },
        [&](int, int, int i, float2 v) { lb[fftc::lpad(i)] = v; });
}

// ---------------------------------------------------------------- columns
// Strided lines (element i of line j at i*es + j): P adjacent lines per
// workgroup, thread = (line-local index, line) with the line fastest, so each
// global access covers P consecutive complex values.  LDS holds [i][line]
// with one pad row after every 16 rows (row lpad(i)): a stage's accesses,
// P lines x 64/P consecutive or 16-strided elements per wave, then spread
// over all LDS banks (a row stride of P + 1 gave 2-way conflicts at
// 8-line panels).
//   Single transforms (MODE 0 / 1): P = 16, a whole 128-byte row segment per
//   access, while the workgroup stays within 1024 threads and 150 KB of LDS
//   (measured at 512-point lines: 25% faster than P = 8).
//   FFT x Qf x IFFT (MODE 2): P = 8 (or fewer for long lines), so two
//   workgroups share a CU (at 1024-point lines P = 16 leaves one
//   1024-thread workgroup per CU and measured 27% slower).
template <int L, int MODE>
constexpr int col_lines() {
    constexpr int T = fftc::threads_per_line(L);
    return (MODE != 2 && T * 16 <= 1024 && (long)L * 17 * 8 <= 150 * 1024)
               ? 16
               : (T * 8 <= 512 ? (T * 16 <= 256 ? 16 : 8) : 512 / T);
}
template <int L, int MODE>
constexpr int col_threads() { return col_lines<L, MODE>() * fftc::threads_per_line(L); }

// MODE 0: forward, 1: inverse, 2: forward, times qf (same addressing with
// batch stride 0), inverse.  HALF_IN: inputs i >= L/2 are zero (the
// zero-padded axis of the 2x grid); HALF_OUT: only outputs i < L/2 are
// stored (the cropped axis).  blockIdx.y selects the batch (bstride).
// RQ (MODE 2): the spectrum is real and read from qr instead of qf.
template <int L, int MODE, bool HALF_IN, bool HALF_OUT, bool RQ = false>
__global__ __launch_bounds__((col_threads<L, MODE>())) void k_fft_cols(
    float2 *data, long bstride, long es, int nlines, const float2 *__restrict__ qf,
    const float2 *__restrict__ tw, const float *__restrict__ qr) {
    constexpr int P = col_lines<L, MODE>();
    constexpr int E = fftc::vals_per_thread(L);
    __shared__ float2 buf[fftc::num_stages(L) > 1 ? fftc::line_lds(L) * P : 1];
    __shared__ float2 twl[tw_in_lds<L>() ? L : 1];
    tw = stage_twiddles<L>(tw, twl);
    const int c = threadIdx.x % P, tl = threadIdx.x / P;
    // Workgroups are dispatched round-robin over the 8 XCDs: blocks b and
    // b + 8 of each group of 16 take adjacent panels, so the two halves of a
    // 128-byte row segment are fetched through the same L2.
    const int bx = blockIdx.x;
    const int panel = (gridDim.x & 15) == 0 ? (bx & ~15) | ((bx & 7) << 1) | ((bx >> 3) & 1) : bx;
    const int j = panel * P + c;
    const bool ok = j < nlines;
    // element i = tl + T m of the line sits at col[m * step]: the register
    // index m is a compile-time constant and step is uniform, so each access
    // is one scalar multiply and one vector add (no per-element 64-bit
    // multiply of i * es)
    constexpr int T = fftc::threads_per_line(L);
    const long step = (long)T * es;
    float2 *col = data + (long)blockIdx.y * bstride + (ok ? j : 0) + (long)tl * es;
    auto lld = [&](int, int, int i) { return buf[fftc::lpad(i) * P + c]; };
    auto lst = [&](int, int, int i, float2 v) { buf[fftc::lpad(i) * P + c] = v; };
    // i < L/2 exactly when m < E/2 (known per register)
    auto gld = [&](int b, int r, int) {
        const int m = fftc::first_reg<L>(b, r);
        const bool in = !HALF_IN || m < E / 2;
        return (ok && in) ? col[m * step] : make_float2(0.f, 0.f);
    };
    auto gst = [&](int b, int r, int, float2 v) {
        const int m = fftc::last_reg<L>(b, r);
        const bool in = !HALF_OUT || m < E / 2;
        if (ok && in) col[m * step] = v;
    };
    if constexpr (MODE == 0) {
        run_fft<L, false>(tl, tw, gld, [] {}, gst, lld, lst);
    } else if constexpr (MODE == 1) {
        run_fft<L, true>(tl, tw, gld, [] {}, gst, lld, lst);
    } else {
        float2 regs[E];
        const long qoff = (ok ? j : 0) + (long)tl * es;
        // Qf_l of this thread's elements, fetched before the forward
        // transform so that their latency hides under it (108 -> 93 us per
        // launch at 1024-point columns)
        using QV = typename std::conditional<RQ, float, float2>::type;
        QV qv[E];
#pragma unroll
        for (int m = 0; m < E; ++m) {
            if constexpr (RQ)
                qv[m] = qr[qoff + m * step];
            else
                qv[m] = qf[qoff + m * step];
        }
        run_fft<L, false>(
            tl, tw, gld, [] {},
            [&](int b, int r, int, float2 v) {
                const int m = fftc::last_reg<L>(b, r);
                if constexpr (RQ)
                    regs[m] = ok ? make_float2(v.x * qv[m], v.y * qv[m]) : v;
                else
                    regs[m] = ok ? cmul(v, qv[m]) : v;
            },
            lld, lst);
        run_fft<L, true>(
            tl, tw, [&](int b, int r, int) { return regs[fftc::first_reg<L>(b, r)]; },
            [] { __syncthreads(); }, gst, lld, lst);
    }
}

// ---------------------------------------------------------------- pass A
// Lines (c, o): image line o (N1*N2o lines of nlast = L/2 elements,
// contiguous), v = S_c * e^{-i 2pi f tau_l} * p, zero-padded to L, forward
// FFT, written to G1 line ((c*N1 + y)*A + x)*L with y = o / N2o, x = o % N2o.
template <int L>
__global__ __launch_bounds__(fftc::row_threads(L)) void k_toep_pass_a(
    const float2 *__restrict__ p, const float2 *__restrict__ sens,
    const float *__restrict__ f, float tau_l, int nimg_lines, int N1, int N2o, int A,
    int ncoils, float2 *__restrict__ g1, const float2 *__restrict__ tw) {
    constexpr int T = fftc::threads_per_line(L);
    constexpr int E = fftc::vals_per_thread(L);
    constexpr int nlast = L / 2;
    __shared__ float2 buf[row_lds<L>()];
    __shared__ float2 twl[tw_in_lds<L>() ? L : 1];
    tw = stage_twiddles<L>(tw, twl);
    const int tl = threadIdx.x % T, ln = threadIdx.x / T;
    const long nvox = (long)nimg_lines * nlast;
    const long line = (long)blockIdx.x * row_lines<L>() + ln;  // flattened (c, o)
    const bool ok = line < (long)ncoils * nimg_lines;
    const int c = ok ? (int)(line / nimg_lines) : 0;
    const long o = ok ? line % nimg_lines : 0;
    const long y = o / N2o, x = o % N2o;
    const float2 *pl = p + o * nlast;
    const float *fl = f + o * nlast;
    const float2 *sl = sens ? sens + (long)c * nvox + o * nlast : nullptr;
    float2 *dst = g1 + (((long)c * N1 + y) * A + x) * L;
    float2 *lb = buf + (fftc::num_stages(L) > 1 ? ln * fftc::line_lds(L) : 0);
    run_fft<L, false>(
        tl, tw,
        [&](int b, int r, int i) {
            if (!ok || fftc::first_reg<L>(b, r) >= E / 2) return make_float2(0.f, 0.f);
            float sn, cs;
            sincos_rev(frac_near(fl[i] * tau_l), &sn, &cs);
            const float2 pv = pl[i];
            // e^{-i 2pi f tau} p
            const float2 u = make_float2(pv.x * cs + pv.y * sn, pv.y * cs - pv.x * sn);
            return sl ? cmul(sl[i], u) : u;
        },
        [] {},
        [&](int, int, int i, float2 v) {
            if (ok) dst[i] = v;
        },
        [&](int, int, int i) { return lb[fftc::lpad(i)]; },
        [&](int, int, int i, float2 v) { lb[fftc::lpad(i)] = v; });
}

// ---------------------------------------------------------------- pass E
// For one image line and the coils of one coil group: inverse FFT of the G1
// lines, keep the first nlast elements, sum conj(S_c) * z_c over the group
// in registers, write partial[group][voxel].
template <int L>
__global__ __launch_bounds__(fftc::row_threads(L)) void k_toep_pass_e(
    const float2 *__restrict__ g1, const float2 *__restrict__ sens, int nimg_lines, int N1,
    int N2o, int A, int ncoils, int coils_per_group, float2 *__restrict__ partial,
    const float2 *__restrict__ tw) {
    constexpr int T = fftc::threads_per_line(L);
    constexpr int E = fftc::vals_per_thread(L);
    constexpr int nlast = L / 2;
    __shared__ float2 buf[row_lds<L>()];
    __shared__ float2 twl[tw_in_lds<L>() ? L : 1];
    tw = stage_twiddles<L>(tw, twl);
    const int tl = threadIdx.x % T, ln = threadIdx.x / T;
    const long nvox = (long)nimg_lines * nlast;
    const long o = (long)blockIdx.x * row_lines<L>() + ln;
    const bool ok = o < nimg_lines;
    const long oc = ok ? o : 0;
    const long y = oc / N2o, x = oc % N2o;
    const int group = blockIdx.y;
    const int c0 = group * coils_per_group;
    const int c1 = min(ncoils, c0 + coils_per_group);
    float2 *lb = buf + (fftc::num_stages(L) > 1 ? ln * fftc::line_lds(L) : 0);
    // only the kept half (i < nlast, i.e. register m < E/2) is accumulated
    float2 acc[E / 2];
#pragma unroll
    for (int m = 0; m < E / 2; ++m) acc[m] = make_float2(0.f, 0.f);
    for (int c = c0; c < c1; ++c) {
        const float2 *src = g1 + (((long)c * N1 + y) * A + x) * L;
        const float2 *sl = sens ? sens + (long)c * nvox + oc * nlast : nullptr;
        run_fft<L, true>(
            tl, tw, [&](int, int, int i) { return ok ? src[i] : make_float2(0.f, 0.f); },
            [] { __syncthreads(); },  // the previous coil's last stage read these lines
            [&](int b, int r, int i, float2 z) {
                const int m = fftc::last_reg<L>(b, r);
                if (m < E / 2) {
                    float2 &a = acc[m];
                    if (sl) {
                        const float2 s = ok ? sl[i] : make_float2(0.f, 0.f);
                        // conj(s) * z
                        a.x += s.x * z.x + s.y * z.y;
                        a.y += s.x * z.y - s.y * z.x;
                    } else {
                        a.x += z.x;
                        a.y += z.y;
                    }
                }
            },
            [&](int, int, int i) { return lb[fftc::lpad(i)]; },
            [&](int, int, int i, float2 v) { lb[fftc::lpad(i)] = v; });
    }
    if (!ok) return;
#pragma unroll
    for (int m = 0; m < E / 2; ++m)
        partial[(long)group * nvox + oc * nlast + tl + T * m] = acc[m];
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
// Q_l in DFT order (index m holds offset m, or m - M past M/2).  Offsets
// of -N on an axis (index M/2) never enter the cropped product (two image
// points differ by at most N - 1 along it), so they are stored as zero:
// Q_l is conjugate-symmetric on all other offsets, and the edge values
// would be the only source of an imaginary part in its spectrum.
__global__ void k_ifftshift_q(const float *__restrict__ qr, const float *__restrict__ qi,
                              int M1, int M2, int M3, float2 *__restrict__ out) {
    const long n = (long)M1 * M2 * M3;
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const int m3 = (int)(m % M3);
    const int m2 = (int)((m / M3) % M2);
    const int m1 = (int)(m / ((long)M3 * M2));
    if (m1 == M1 / 2 || m2 == M2 / 2 || (M3 > 1 && m3 == M3 / 2)) {
        out[m] = make_float2(0.f, 0.f);
        return;
    }
    const int s1 = (m1 + M1 / 2) % M1;
    const int s2 = (m2 + M2 / 2) % M2;
    const int s3 = M3 > 1 ? (m3 + M3 / 2) % M3 : 0;
    const long src = ((long)s1 * M2 + s2) * M3 + s3;
    out[m] = make_float2(qr[src], qi[src]);
}

// Sets *asym when Q_l (DFT order) is not exactly conjugate-symmetric:
// q[m] != conj(q[-m]) for some m (indices mod M per axis).
__global__ void k_q_conj_check(const float2 *__restrict__ q, int M1, int M2, int M3,
                               unsigned int *asym) {
    const long n = (long)M1 * M2 * M3;
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const int m3 = (int)(m % M3);
    const int m2 = (int)((m / M3) % M2);
    const int m1 = (int)(m / ((long)M3 * M2));
    const long pm = ((long)((M1 - m1) % M1) * M2 + (M2 - m2) % M2) * M3 + (M3 - m3) % M3;
    const float2 a = q[m], b = q[pm];
    if (a.x != b.x || a.y != -b.y) atomicOr(asym, 1u);
}

__global__ void k_zero_flag(unsigned int *p) { *p = 0u; }

__global__ void k_qf_real(const float2 *__restrict__ qf, long n, float *__restrict__ qr) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        qr[i] = qf[i].x;
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
        case 24: { constexpr int L_ = 24; BODY; } break;     \
        case 48: { constexpr int L_ = 48; BODY; } break;     \
        case 96: { constexpr int L_ = 96; BODY; } break;     \
        case 192: { constexpr int L_ = 192; BODY; } break;   \
        case 384: { constexpr int L_ = 384; BODY; } break;   \
        case 768: { constexpr int L_ = 768; BODY; } break;   \
        case 1536: { constexpr int L_ = 1536; BODY; } break; \
        default: return fail(IMP_EUNSUPPORTED, "FFT length %d unsupported", Lv); \
    }

// lengths with an instantiated kernel set: 2^k (2..2048), 3 * 2^k (24..1536)
bool fft_len_ok(int L) {
    return fftc::fft_length_supported(L) && L <= 2048 && (L % 3 != 0 || L >= 24);
}

}  // namespace

// ------------------------------------------------------------ twiddles
int twiddle_table_create(imp_ctx *ctx) {
    // table for length L: powers of two at offset L, lengths 3 * 2^k at
    // offset 4096 + L (8192 entries, FP64 values rounded once)
    if (ctx->twiddle) return IMP_OK;
    std::vector<float2> h(8192, make_float2(0.f, 0.f));
    auto fill = [&](int n, int off) {
        for (int m = 0; m < n; ++m) {
            const double a = -2.0 * M_PI * (double)m / (double)n;
            h[off + m] = make_float2((float)cos(a), (float)sin(a));
        }
    };
    for (int n = 2; n <= 2048; n *= 2) fill(n, n);
    for (int n = 24; n <= 1536; n *= 2) fill(n, 4096 + n);
    IMP_HIP(hipMalloc(&ctx->twiddle, 8192 * sizeof(float2)));
    IMP_HIP(hipMemcpyAsync(ctx->twiddle, h.data(), 8192 * sizeof(float2), hipMemcpyHostToDevice,
                           ctx->stream));
    IMP_HIP(hipStreamSynchronize(ctx->stream));
    return IMP_OK;
}

int twiddles(imp_ctx *ctx, int L, const float2 **out) {
    if (!fft_len_ok(L)) return fail(IMP_EUNSUPPORTED, "FFT length %d unsupported", L);
    IMP_TRY(twiddle_table_create(ctx));
    *out = ctx->twiddle + (L % 3 == 0 ? 4096 + L : L);
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
        const int NL = row_lines<L_>();
        const dim3 block(fftc::row_threads(L_));
        if (inverse)
            hipLaunchKernelGGL((k_fft_rows<L_, true>), dim3(ceil_div(nl, NL)), block, 0,
                               ctx->stream, data, nl, tw);
        else
            hipLaunchKernelGGL((k_fft_rows<L_, false>), dim3(ceil_div(nl, NL)), block, 0,
                               ctx->stream, data, nl, tw);
    });
    IMP_CHECK_LAUNCH();
    if (d3) {  // x axis: batch y, lines z
        IMP_TRY(twiddles(ctx, M2, &tw));
        IMP_L_SWITCH(M2, {
            dim3 grid(ceil_div(M3, col_lines<L_, 0>()), M1);
            if (inverse)
                hipLaunchKernelGGL((k_fft_cols<L_, 1, false, false>), grid,
                                   dim3(col_threads<L_, 1>()), 0, ctx->stream, data, (long)M2 * M3,
                                   (long)M3, M3, nullptr, tw, nullptr);
            else
                hipLaunchKernelGGL((k_fft_cols<L_, 0, false, false>), grid,
                                   dim3(col_threads<L_, 0>()), 0, ctx->stream, data, (long)M2 * M3,
                                   (long)M3, M3, nullptr, tw, nullptr);
        });
        IMP_CHECK_LAUNCH();
    }
    // y axis: one batch, lines = M2*M3
    IMP_TRY(twiddles(ctx, M1, &tw));
    const int nlines = M2 * M3;
    IMP_L_SWITCH(M1, {
        dim3 grid(ceil_div(nlines, col_lines<L_, 0>()), 1);
        if (inverse)
            hipLaunchKernelGGL((k_fft_cols<L_, 1, false, false>), grid, dim3(col_threads<L_, 1>()),
                               0, ctx->stream, data, 0L, (long)nlines, nlines, nullptr, tw, nullptr);
        else
            hipLaunchKernelGGL((k_fft_cols<L_, 0, false, false>), grid, dim3(col_threads<L_, 0>()),
                               0, ctx->stream, data, 0L, (long)nlines, nlines, nullptr, tw, nullptr);
    });
    IMP_CHECK_LAUNCH();
    return IMP_OK;
}

// Qf_l = FFT(ifftshift(Q_l)) for all l, into plan-owned HBM.
int toep_prepare_qf(imp_ctx *ctx, ToepPlan &P, const float *q_r, const float *q_i) {
    const long G = P.G;
    // With the edges zeroed, a Q_l that is exactly conjugate-symmetric
    // (Q_l(-r) = conj(Q_l(r)) bit for bit, as the standard-grid kernels
    // write it) has a real spectrum: its imaginary part is FFT rounding, and
    // pass C multiplies by the real plane (half the bytes per element).
    // Any other Q_l (read from files, computed point by point on explicit
    // grids) keeps the complex spectrum as given.
    IMP_TRY(ensure_dscratch(ctx, 1));
    unsigned int *asym = (unsigned int *)ctx->dscratch;
    hipLaunchKernelGGL(k_zero_flag, dim3(1), dim3(1), 0, ctx->stream, asym);
    for (int l = 0; l < P.nseg; ++l) {
        float2 *dst = P.qf + (long)l * G;
        hipLaunchKernelGGL(k_ifftshift_q, dim3(ceil_div(G, 256)), dim3(256), 0, ctx->stream,
                           q_r + (long)l * G, q_i + (long)l * G, P.M1, P.M2, P.M3, dst);
        hipLaunchKernelGGL(k_q_conj_check, dim3(ceil_div(G, 256)), dim3(256), 0, ctx->stream,
                           dst, P.M1, P.M2, P.M3, asym);
        IMP_CHECK_LAUNCH();
        IMP_TRY(fft_grid(ctx, dst, P.M1, P.M2, P.M3, false));
    }
    unsigned int h = 1u;
    IMP_HIP(hipMemcpyAsync(&h, asym, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    IMP_HIP(hipStreamSynchronize(ctx->stream));
    P.qf_real = h == 0u && P.qfr;
    if (P.qf_real) {
        hipLaunchKernelGGL(k_qf_real, dim3(1024), dim3(256), 0, ctx->stream, P.qf,
                           (long)P.nseg * G, P.qfr);
        IMP_CHECK_LAUNCH();
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
    if (2 * nlast != Ml)
        return fail(IMP_EINVAL, "Toeplitz apply: last axis %d is not 2 x %d", Ml, nlast);
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
            const long nlines = (long)C * nimg_lines;
            hipLaunchKernelGGL((k_toep_pass_a<L_>), dim3(ceil_div(nlines, row_lines<L_>())),
                               dim3(fftc::row_threads(L_)), 0, ctx->stream, p, P.sens, P.fcyc, tau_l,
                               nimg_lines, P.N1, N2o, A, C, P.g1, tw_last);
        });
        prof_end(ctx, rec);
        IMP_CHECK_LAUNCH();
        // pass B (3D): x axis, batches (c, y), lines z in [0, M3), input x < N2
        if (d3) {
            rec = prof_begin(ctx, "toep_pass_b");
            IMP_L_SWITCH(P.M2, {
                dim3 grid(ceil_div(P.M3, col_lines<L_, 0>()), C * P.N1);
                hipLaunchKernelGGL((k_fft_cols<L_, 0, true, false>), grid, dim3(col_threads<L_, 0>()),
                                   0, ctx->stream, P.g1, (long)P.M2 * P.M3, (long)P.M3, P.M3,
                                   nullptr, tw_x, nullptr);
            });
            prof_end(ctx, rec);
            IMP_CHECK_LAUNCH();
        }
        // pass C: y axis with Qf multiply, batches c, lines (x,z) in [0, A*Ml)
        {
            const int nl = A * Ml;
            const float2 *qf = P.qf + (long)l * P.G;
            rec = prof_begin(ctx, "toep_pass_c");
            const float *qr = P.qf_real ? P.qfr + (long)l * P.G : nullptr;
            IMP_L_SWITCH(P.M1, {
                dim3 grid(ceil_div(nl, col_lines<L_, 2>()), C);
                if (qr)
                    hipLaunchKernelGGL((k_fft_cols<L_, 2, true, true, true>), grid,
                                       dim3(col_threads<L_, 2>()), 0, ctx->stream, P.g1,
                                       (long)P.N1 * nl, (long)nl, nl, nullptr, tw_y, qr);
                else
                    hipLaunchKernelGGL((k_fft_cols<L_, 2, true, true>), grid,
                                       dim3(col_threads<L_, 2>()), 0, ctx->stream, P.g1,
                                       (long)P.N1 * nl, (long)nl, nl, qf, tw_y, nullptr);
            });
            prof_end(ctx, rec);
            IMP_CHECK_LAUNCH();
        }
        // pass D (3D): inverse x, keep x < N2
        if (d3) {
            rec = prof_begin(ctx, "toep_pass_d");
            IMP_L_SWITCH(P.M2, {
                dim3 grid(ceil_div(P.M3, col_lines<L_, 1>()), C * P.N1);
                hipLaunchKernelGGL((k_fft_cols<L_, 1, false, true>), grid, dim3(col_threads<L_, 1>()),
                                   0, ctx->stream, P.g1, (long)P.M2 * P.M3, (long)P.M3, P.M3,
                                   nullptr, tw_x, nullptr);
            });
            prof_end(ctx, rec);
            IMP_CHECK_LAUNCH();
        }
        // pass E: inverse last axis, conj(S_c), partial sums over coil groups
        rec = prof_begin(ctx, "toep_pass_e");
        IMP_L_SWITCH(Ml, {
            dim3 grid(ceil_div(nimg_lines, row_lines<L_>()), ngroups);
            hipLaunchKernelGGL((k_toep_pass_e<L_>), grid, dim3(fftc::row_threads(L_)), 0,
                               ctx->stream, P.g1, P.sens, nimg_lines, P.N1, N2o, A, C, cpg,
                               P.partial, tw_last);
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

// The 2x grid (Toeplitz passes) and the image grid (preconditioner) both
// need instantiated FFT lengths.
bool toep_sizes_supported(int N1, int N2, int N3) {
    if (!fft_len_ok(2 * N1) || !fft_len_ok(2 * N2)) return false;
    if (!fft_len_ok(N1) || !fft_len_ok(N2)) return false;
    if (N3 > 1 && (!fft_len_ok(2 * N3) || !fft_len_ok(N3))) return false;
    return true;
}

}  // namespace imp
