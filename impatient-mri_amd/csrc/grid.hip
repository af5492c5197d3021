// Gridding path of the Toeplitz method (-toeplitzGridding): Q_l and F^H d
// by Kaiser-Bessel gridding onto an oversampled Cartesian grid and a shifted,
// centre-cropped inverse DFT, as computeQ_GPU_Grid / computeFH_GPU_Grid do
// (toeplitz/computeQ.cmem.cu:401-754, computeFH.cmem.cu:470-802,
// 1032-1117; binning and gather in CUDA_interface.cu:26-400 and
// GPU_kernels.cu:68-581; crop, deapodization and shifts in
// gridding_utils.cu:6-290).
//
// The estimate per transform, for samples k_j (cycles of an n-point image
// axis) with weights v_j:
//   1. s_j = os (k_j + n/2) per axis; grid point g in [0, G) receives
//      v_j prod_axes KB(|s_j - g| / os) for |s_j - g| / os <= W/2 (W = 4),
//      KB(d) = I0(beta sqrt(1 - (2d/W)^2)) / W; the grid does not wrap;
//   2. out(i) = sum_g grid(g) e^{+i 2pi (g - G/2)(i - n/2) / G} per axis
//      (ifftshift, unnormalised inverse FFT, fftshift, centre crop);
//   3. out(i) /= os C(u_i) per axis, C the kernel's transform at
//      u_i = (i - n/2)/n.
//
// This build's structure:
//   - samples are sorted by grid cell (a stable radix sort, so each cell's
//     samples stay in index order), so each grid point gathers from
//     contiguous sample ranges with no atomics on the grid;
//   - one gather launch grids every coil of a segment at once (each sample's
//     kernel weight serves all coils);
//   - the DFT of step 2 runs per axis on lines of length G = R P (P the
//     largest power of two dividing G): R radix-2 FFTs of length P in LDS,
//     then the R-term combination evaluated only at the n kept outputs, with
//     step 3 folded in; every pass writes its output with the transformed
//     axis outermost, so after one pass per axis the layout is the image's.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"
#include "toeplitz.h"

namespace imp {

namespace {

constexpr float KB_WIDTH = 4.0f;
constexpr int GT = 256;

// I0(x), x >= 0, by the polynomial approximations of Abramowitz & Stegun,
// Handbook of Mathematical Functions, 9.8.1 (|err| < 1.6e-7 for x <= 3.75)
// and 9.8.2 (relative error < 1.9e-7 above).
__device__ __forceinline__ float bessel_i0(float x) {
    if (x < 3.75f) {
        float y = x * (1.0f / 3.75f);
        y *= y;
        return 1.0f + y * (3.5156229f + y * (3.0899424f + y * (1.2067492f + y * (0.2659732f +
                     y * (0.0360768f + y * 0.0045813f)))));
    }
    const float y = 3.75f / x;
    const float p = 0.39894228f + y * (0.01328592f + y * (0.00225319f + y * (-0.00157565f +
                    y * (0.00916281f + y * (-0.02057706f + y * (0.02635537f +
                    y * (-0.01647633f + y * 0.00392377f)))))));
    return p * expf(x) * rsqrtf(x);
}

// KB(d) = I0(beta sqrt(1 - (2d/W)^2)) / W for d = |s - g| / os <= W/2, else 0
__device__ __forceinline__ float kb_weight(float s, int g, float inv_os, float beta) {
    const float d = fabsf(s - (float)g) * inv_os;
    if (!(d <= 0.5f * KB_WIDTH)) return 0.f;
    const float u = 2.0f * d * (1.0f / KB_WIDTH);
    return bessel_i0(beta * sqrtf(fmaxf(1.0f - u * u, 0.0f))) * (1.0f / KB_WIDTH);
}

struct GridGeom {
    int G[3];        // grid size per axis (x, y, z); G[2] = 1 in 2D
    int n[3];        // image size per axis
    float os, inv_os, half, beta;
    int d3;
};

// shifted positions s = os (k + n/2) (float, as the reference computes them),
// and the sort key of every sample: its time bucket b (segment l weighs only
// buckets l and l + 1; b = floor((t - t0)/tau) + 1 clamped to [0, nseg]) and
// its grid cell ((cy Gx) + cx) Gz + cz, as b ncells + cell
__global__ void k_grid_cells(int m, const float *kx, const float *ky, const float *kz,
                             const float *t, float kscale, GridGeom g, float t0, float tau,
                             int nseg, float4 *pos, int *key, int *count) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const float sx = g.os * (kscale * kx[j] + (float)g.n[0] / 2.0f);
    const float sy = g.os * (kscale * ky[j] + (float)g.n[1] / 2.0f);
    const float sz = g.d3 ? g.os * (kscale * kz[j] + (float)g.n[2] / 2.0f) : 0.f;
    const int cx = min(g.G[0] - 1, max(0, (int)floorf(sx)));
    const int cy = min(g.G[1] - 1, max(0, (int)floorf(sy)));
    const int cz = g.d3 ? min(g.G[2] - 1, max(0, (int)floorf(sz))) : 0;
    const int c = ((cy * g.G[0]) + cx) * g.G[2] + cz;
    int b = 0;
    if (tau > 0.f) {
        const double f = floor(((double)t[j] - (double)t0) / (double)tau);
        b = (int)fmin(fmax(f + 1.0, 0.0), (double)nseg);
    }
    const long ncells = (long)g.G[0] * g.G[1] * g.G[2];
    const int k = (int)(b * ncells + c);
    pos[j] = make_float4(sx, sy, sz, 0.f);
    key[j] = k;
    atomicAdd(&count[k], 1);
}

// exclusive scan of n ints in three launches: per-block scans with block
// totals, a scan of the totals by one workgroup, then the block offsets
__global__ __launch_bounds__(GT) void k_scan_blocks(const int *in, int n, int *out,
                                                    int *block_sum) {
    __shared__ int sm[GT];
    const int base = blockIdx.x * GT * 4;
    int v[4], s = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int i = base + threadIdx.x * 4 + e;
        v[e] = i < n ? in[i] : 0;
        s += v[e];
    }
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < GT; off <<= 1) {
        const int t = threadIdx.x >= off ? sm[threadIdx.x - off] : 0;
        __syncthreads();
        sm[threadIdx.x] += t;
        __syncthreads();
    }
    int run = sm[threadIdx.x] - s;   // exclusive prefix of this thread
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int i = base + threadIdx.x * 4 + e;
        if (i < n) out[i] = run;
        run += v[e];
    }
    if (threadIdx.x == GT - 1) block_sum[blockIdx.x] = sm[GT - 1];
}

__global__ __launch_bounds__(GT) void k_scan_totals(int *block_sum, int nb, int *total) {
    __shared__ int sm[GT];
    int carry = 0;
    for (int b0 = 0; b0 < nb; b0 += GT) {
        const int i = b0 + threadIdx.x;
        const int v = i < nb ? block_sum[i] : 0;
        sm[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < GT; off <<= 1) {
            const int t = threadIdx.x >= off ? sm[threadIdx.x - off] : 0;
            __syncthreads();
            sm[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) block_sum[i] = carry + sm[threadIdx.x] - v;
        const int tot = sm[GT - 1];
        __syncthreads();
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void k_scan_add(int *out, int n, const int *block_sum, const int *total) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += block_sum[i / (GT * 4)];
    if (i == n) out[n] = *total;
}

// Stable sort of the samples by cell (the gather then sums each cell's
// samples in index order, so every run gives the same bits): LSD radix sort
// of (cell, sample index) pairs, 4-bit digits, one element per thread and
// RS_BLOCK elements per workgroup.  Per pass:
//   k_rs_count    digit histogram of each block, stored digit-major
//                 hist[d * nblocks + block];
//   (exclusive scan of hist: every (digit, block) gets its output offset;
//   digits ascending, blocks in order)
//   k_rs_scatter  each element goes to its (digit, block) offset plus the
//                 number of earlier elements of its block with that digit
//                 (wave ballots within a wave, per-wave counts in LDS
//                 across waves), which keeps equal keys in input order.
constexpr int RS_BLOCK = 256;
constexpr int RS_WAVES = RS_BLOCK / 64;
constexpr int RS_BINS = 16;

__global__ __launch_bounds__(RS_BLOCK) void k_rs_count(const int *__restrict__ key, int m,
                                                       int shift, int nblocks,
                                                       int *__restrict__ hist) {
    __shared__ int cnt[RS_BINS];
    if (threadIdx.x < RS_BINS) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int i = blockIdx.x * RS_BLOCK + threadIdx.x;
    if (i < m) atomicAdd(&cnt[(key[i] >> shift) & (RS_BINS - 1)], 1);
    __syncthreads();
    if (threadIdx.x < RS_BINS) hist[threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

__global__ __launch_bounds__(RS_BLOCK) void k_rs_scatter(const int *__restrict__ key_in,
                                                         const int *__restrict__ val_in, int m,
                                                         int shift, int nblocks,
                                                         const int *__restrict__ offs,
                                                         int *__restrict__ key_out,
                                                         int *__restrict__ val_out) {
    __shared__ int wcnt[RS_WAVES][RS_BINS];
    const int i = blockIdx.x * RS_BLOCK + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool ok = i < m;
    const int k = ok ? key_in[i] : 0;
    const int d = ok ? (k >> shift) & (RS_BINS - 1) : RS_BINS;   // RS_BINS: no element
    const unsigned long long lt = (1ull << lane) - 1ull;
    int rank = 0;
#pragma unroll
    for (int bin = 0; bin < RS_BINS; ++bin) {
        const unsigned long long mk = __ballot(d == bin);
        if (d == bin) rank = __popcll(mk & lt);
        if (lane == 0) wcnt[wave][bin] = __popcll(mk);
    }
    __syncthreads();
    if (!ok) return;
    int before = 0;
    for (int w = 0; w < wave; ++w) before += wcnt[w][d];
    const int dst = offs[d * nblocks + blockIdx.x] + before + rank;
    key_out[dst] = k;
    val_out[dst] = val_in[i];
}

__global__ void k_iota(int m, int *v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) v[i] = i;
}

// segment weights of every sample: a_l(t) is nonzero for at most two
// consecutive l; w.x = a_{l0}, w.y = a_{l0+1}, l0 in pos.w (as int bits).
// hanning_d of computeQ.cmem.cu:125-138 / GPU_kernels.cu:48-62 in FP64.
__global__ void k_grid_samples(int m, const int *perm, const float4 *pos_in, const float *t,
                               int nseg, float tau, float t0, float4 *pos, float2 *w) {
    const int ks = blockIdx.x * blockDim.x + threadIdx.x;
    if (ks >= m) return;
    const int j = perm[ks];
    float4 p = pos_in[j];
    const float tm = t[j];
    float wl[2] = {0.f, 0.f};
    int l0 = 0;
    if (tau > 0.f) {   // tau = 0 (one readout time): hanning_d is 0 everywhere
        l0 = (int)floor(((double)tm - (double)t0) / (double)tau);
        for (int q = 0; q < 2; ++q) {
            const int l = l0 + q;
            if (l < 0 || l >= nseg) continue;
            const float taul = tau * (float)l;
            const float x = tm - taul - t0;
            if (fabsf(x) < tau) wl[q] = (float)(0.5 + 0.5 * cospi((double)x / (double)tau));
        }
    }
    p.w = __int_as_float(l0);
    pos[ks] = p;
    w[ks] = make_float2(wl[0], wl[1]);
}

// coil data in sorted sample order, [ks][C] (unit data when d_r is null)
template <int C>
__global__ void k_grid_data(int m, int ncoils, const int *perm, const float *d_r,
                            const float *d_i, float2 *out) {
    const int ks = blockIdx.x * blockDim.x + threadIdx.x;
    if (ks >= m) return;
    const int j = perm[ks];
#pragma unroll
    for (int c = 0; c < C; ++c)
        out[(size_t)ks * C + c] = c < ncoils ? make_float2(d_r[(size_t)c * m + j],
                                                          d_i[(size_t)c * m + j])
                                             : make_float2(0.f, 0.f);
}

// The window of grid point id ([gy][gx][gz], the innermost cell axis
// fastest): the cells whose samples can lie within W/2 of it.
struct Window {
    int gx, gy, gz, x0, x1, y0, y1, z0, z1;
};

__device__ __forceinline__ Window window_of(const GridGeom &g, long id) {
    Window W;
    if (g.d3) {
        W.gz = (int)(id % g.G[2]);
        W.gx = (int)((id / g.G[2]) % g.G[0]);
        W.gy = (int)(id / ((long)g.G[2] * g.G[0]));
    } else {
        W.gz = 0;
        W.gx = (int)(id % g.G[0]);
        W.gy = (int)(id / g.G[0]);
    }
    const int r_lo = (int)ceilf(g.half), r_hi = (int)floorf(g.half);
    W.y0 = max(0, W.gy - r_lo);
    W.y1 = min(g.G[1] - 1, W.gy + r_hi);
    W.x0 = max(0, W.gx - r_lo);
    W.x1 = min(g.G[0] - 1, W.gx + r_hi);
    W.z0 = g.d3 ? max(0, W.gz - r_lo) : 0;
    W.z1 = g.d3 ? min(g.G[2] - 1, W.gz + r_hi) : 0;
    return W;
}

// f(b, e) for each contiguous sample range [b, e) of the window: per time
// bucket l and l + 1 of segment l, per cell row (y, x) in 3D with z
// contiguous, per row y in 2D with x contiguous
template <class F>
__device__ __forceinline__ void for_ranges(const GridGeom &g, const int *start, int nbuck, int l,
                                           const Window &W, F f) {
    const long total = (long)g.G[0] * g.G[1] * g.G[2];
    const int o1 = g.d3 ? W.x0 : 0, o1e = g.d3 ? W.x1 : 0;
    const int b1 = min(l + 1, nbuck - 1);
    for (int bk = l; bk <= b1; ++bk) {
        const long bbase = (long)bk * total;
        for (int cy = W.y0; cy <= W.y1; ++cy) {
            for (int cx = o1; cx <= o1e; ++cx) {
                if (g.d3) {
                    const long row = bbase + ((long)cy * g.G[0] + cx) * g.G[2];
                    f(start[row + W.z0], start[row + W.z1 + 1]);
                } else {
                    const long row = bbase + (long)cy * g.G[0];
                    f(start[row + W.x0], start[row + W.x1 + 1]);
                }
            }
        }
    }
}

// samples b + off, b + off + stride, ... < e of segment l into the point's sums
template <int C>
__device__ __forceinline__ void accumulate(const GridGeom &g, const Window &W,
                                           const float4 *__restrict__ pos,
                                           const float2 *__restrict__ w,
                                           const float2 *__restrict__ data, int l, int b, int e,
                                           int off, int stride, float *acc_r, float *acc_i) {
    constexpr int NC = C > 0 ? C : 1;
    for (int ks = b + off; ks < e; ks += stride) {
        const float4 p = pos[ks];
        const float2 ww = w[ks];
        const int l0 = __float_as_int(p.w);
        const float a = l == l0 ? ww.x : (l == l0 + 1 ? ww.y : 0.f);
        if (a == 0.f) continue;
        float kw = kb_weight(p.x, W.gx, g.inv_os, g.beta);
        if (kw == 0.f) continue;
        kw *= kb_weight(p.y, W.gy, g.inv_os, g.beta);
        if (g.d3) kw *= kb_weight(p.z, W.gz, g.inv_os, g.beta);
        const float s = kw * a;
        if (C == 0) {
            acc_r[0] += s;
        } else {
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const float2 v = data[(size_t)ks * NC + c];
                acc_r[c] = fmaf(s, v.x, acc_r[c]);
                acc_i[c] = fmaf(s, v.y, acc_i[c]);
            }
        }
    }
}

// a point whose window holds more samples than this is summed by a whole
// wave (k_grid_gather_heavy): near the centre of a variable-density
// trajectory a window holds thousands of samples, and one lane per point
// would leave its wave waiting on that lane
constexpr int GATHER_HEAVY = 64;

// One thread per grid point: sum over the samples within the kernel window
// of segment l (weight a_l), for C coils (C = 0: unit data, one plane).
// Grid layout per coil: [gy][gx][gz], written at grid + c * plane.  Points
// with more than GATHER_HEAVY samples are listed in heavy[] (in any order)
// and left to k_grid_gather_heavy.
template <int C>
__global__ __launch_bounds__(GT) void k_grid_gather(GridGeom g, const int *start, int nbuck,
                                                    const float4 *pos, const float2 *w,
                                                    const float2 *data, int l, float2 *grid,
                                                    size_t plane, int *heavy, int *nheavy) {
    constexpr int NC = C > 0 ? C : 1;
    const long id = (long)blockIdx.x * GT + threadIdx.x;
    const long total = (long)g.G[0] * g.G[1] * g.G[2];
    if (id >= total) return;
    const Window W = window_of(g, id);
    int count = 0;
    for_ranges(g, start, nbuck, l, W, [&](int b, int e) { count += e - b; });
    if (count > GATHER_HEAVY) {
        heavy[atomicAdd(nheavy, 1)] = (int)id;
        return;
    }
    float acc_r[NC], acc_i[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc_r[c] = acc_i[c] = 0.f;
    for_ranges(g, start, nbuck, l, W, [&](int b, int e) {
        accumulate<C>(g, W, pos, w, data, l, b, e, 0, 1, acc_r, acc_i);
    });
#pragma unroll
    for (int c = 0; c < NC; ++c) grid[c * plane + id] = make_float2(acc_r[c], acc_i[c]);
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// The listed points, one wave each: lane k sums the window's samples
// k, k + 64, ... of every range, then a fixed butterfly adds the lanes, so
// the result does not depend on which wave takes the point.  Waves stride
// over the list; every wave reads the same count and ends.
template <int C>
__global__ __launch_bounds__(GT) void k_grid_gather_heavy(GridGeom g, const int *start, int nbuck,
                                                          const float4 *pos, const float2 *w,
                                                          const float2 *data, int l,
                                                          float2 *grid, size_t plane,
                                                          const int *heavy,
                                                          const int *nheavy) {
    constexpr int NC = C > 0 ? C : 1;
    const int lane = threadIdx.x & 63;
    const int wave = (int)((blockIdx.x * GT + threadIdx.x) >> 6);
    const int nwaves = (int)(gridDim.x * (GT / 64));
    const int n = *nheavy;
    for (int it = wave; it < n; it += nwaves) {
        const long id = heavy[it];
        const Window W = window_of(g, id);
        float acc_r[NC], acc_i[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) acc_r[c] = acc_i[c] = 0.f;
        for_ranges(g, start, nbuck, l, W, [&](int b, int e) {
            accumulate<C>(g, W, pos, w, data, l, b, e, lane, 64, acc_r, acc_i);
        });
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            acc_r[c] = wave_sum(acc_r[c]);
            acc_i[c] = wave_sum(acc_i[c]);
        }
        if (lane == 0) {
#pragma unroll
            for (int c = 0; c < NC; ++c) grid[c * plane + id] = make_float2(acc_r[c], acc_i[c]);
        }
    }
}

// Cropped, shifted inverse DFT along the contiguous axis of every line:
//   out[b][i][a] = dp[i] sum_{g<G} in[b][a][g] e^{+i 2pi g (i - n/2) / G}
// (dp carries (-1)^(i - n/2) of the shift and 1 / (os C(u_i))), for lines
// a < A of batch item b.  G = R P: R radix-2 FFTs of length P in LDS over
// the subsequences g = R q + j, then the R-term combination at each kept i,
//   out(nn) = sum_j Y_j[nn mod P] e^{+i 2pi j nn / G},  nn = i - n/2,
// with the twiddle index j nn mod G stepped by additions.  tw[k] =
// e^{+i 2pi k / G}.  2^logLPW lines per workgroup; every index is 32-bit
// and every division by a power of two is a shift (q = g / R by a float
// reciprocal and one correction step).
__global__ __launch_bounds__(GT) void k_line_dft(const float2 *__restrict__ in,
                                                 float2 *__restrict__ out, int G, int R, int P,
                                                 int logP, int n, int A, int nlines, int logLPW,
                                                 float rcpR, const float2 *__restrict__ tw,
                                                 const float *__restrict__ dp) {
    extern __shared__ float2 lds[];
    const int LPW = 1 << logLPW;
    const int line0 = blockIdx.x * LPW;
    const int nl = min(LPW, nlines - line0);
    // load: each line in natural order (coalesced), scattered to the
    // bit-reversed position of its subsequence
    for (int li = 0; li < nl; ++li) {
        const float2 *__restrict__ src = in + (size_t)(line0 + li) * G;
        float2 *dst = lds + li * G;
        for (int g = threadIdx.x; g < G; g += GT) {
            int q = (int)((float)g * rcpR);
            int j = g - q * R;
            if (j < 0) {
                --q;
                j += R;
            } else if (j >= R) {
                ++q;
                j -= R;
            }
            const int qr = logP ? (int)(__brev((unsigned)q) >> (32 - logP)) : 0;
            dst[j * P + qr] = src[g];
        }
    }
    __syncthreads();
    // radix-2 stages: butterfly span h = 2^lh = 1, 2, ..., P/2
    if (logP > 0) {
        const int nb = (nl * R) << (logP - 1);
        for (int lh = 0; lh < logP; ++lh) {
            const int h = 1 << lh;
            const int tstride = G >> (lh + 1);                 // G / (2h), exact
            const int gmask = (P >> (lh + 1)) - 1;             // P / (2h) - 1
            for (int bi = threadIdx.x; bi < nb; bi += GT) {
                const int k = bi & (h - 1);
                const int grp = (bi >> lh) & gmask;
                const int sub = bi >> (logP - 1);              // (line, j)
                const int base = sub * P + grp * 2 * h + k;
                const float2 wv = tw[k * tstride];
                const float2 u = lds[base], v = lds[base + h];
                const float2 tv = make_float2(v.x * wv.x - v.y * wv.y, v.x * wv.y + v.y * wv.x);
                lds[base] = make_float2(u.x + tv.x, u.y + tv.y);
                lds[base + h] = make_float2(u.x - tv.x, u.y - tv.y);
            }
            __syncthreads();
        }
    }
    // combination at the kept outputs, written with the axis outermost:
    // consecutive threads take consecutive lines of one output index
    const int b0 = line0 / A, a0 = line0 - b0 * A;
    for (int e = threadIdx.x; e < (n << logLPW); e += GT) {
        const int li = e & (LPW - 1), i = e >> logLPW;
        if (li >= nl) continue;
        const int nn = i - n / 2;
        const int m = nn & (P - 1);                            // nn mod P (P a power of two)
        int step = nn % G;
        if (step < 0) step += G;
        const float2 *y = lds + li * G + m;
        float sr = 0.f, si = 0.f;
        int kk = 0;
        for (int j = 0; j < R; ++j) {
            const float2 yv = y[j * P];
            const float2 wv = tw[kk];
            sr += yv.x * wv.x - yv.y * wv.y;
            si += yv.x * wv.y + yv.y * wv.x;
            kk += step;
            if (kk >= G) kk -= G;
        }
        int a = a0 + li, b = b0;
        while (a >= A) {
            a -= A;
            ++b;
        }
        const float f = dp[i];
        out[((size_t)b * n + i) * A + a] = make_float2(sr * f, si * f);
    }
}

__global__ void k_twiddle_g(int G, float2 *tw) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < G) {
        double s, c;
        sincospi(2.0 * (double)k / (double)G, &s, &c);
        tw[k] = make_float2((float)c, (float)s);
    }
}

// dp[i] = (-1)^(i - n/2) / (os C(u_i)), C(u) = sin(sqrt(pi^2 W^2 u^2 -
// beta^2)) / sqrt(...) (sinh below zero), u_i = (i - n/2) / n, in FP64
__global__ void k_deapod(int n, double os, double beta, float *dp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int nn = i - n / 2;
    const double u = (double)nn / (double)n;
    const double e = M_PI * M_PI * (double)KB_WIDTH * (double)KB_WIDTH * u * u - beta * beta;
    double c;
    if (e > 0.0) {
        const double r = sqrt(e);
        c = sin(r) / r;
    } else if (e < 0.0) {
        const double r = sqrt(-e);
        c = sinh(r) / r;
    } else {
        c = 1.0;
    }
    dp[i] = (float)(((nn & 1) ? -1.0 : 1.0) / (os * c));
}

// F^H d epilogue of segment l: acc += e^{+i w (t0 + l tau)} sum_c conj(S_c) y_c
__global__ void k_grid_fhd_accum(long nv, int ncoils, const float2 *y, size_t plane,
                                 const float *s_r, const float *s_i, const float *fm,
                                 float tl, double2 *acc) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    double vr = 0.0, vi = 0.0;
    for (int c = 0; c < ncoils; ++c) {
        const float2 v = y[c * plane + i];
        const double sr = s_r ? s_r[(size_t)c * nv + i] : 1.0;
        const double si = s_r ? s_i[(size_t)c * nv + i] : 0.0;
        vr += sr * v.x + si * v.y;
        vi += sr * v.y - si * v.x;
    }
    // the reference's expfm: cos / sin of the float product fm (tau l + t0)
    const float arg = fm ? fm[i] * tl : 0.f;
    double sn, cs;
    sincos((double)arg, &sn, &cs);
    double2 a = acc[i];
    a.x += cs * vr - sn * vi;
    a.y += cs * vi + sn * vr;
    acc[i] = a;
}

__global__ void k_grid_fhd_out(long nv, const double2 *acc, float *out_r, float *out_i) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nv) {
        out_r[i] = (float)acc[i].x;
        out_i[i] = (float)acc[i].y;
    }
}

__global__ void k_grid_q_out(long nq, const float2 *y, float *q_r, float *q_i) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nq) {
        q_r[i] = y[i].x;
        q_i[i] = y[i].y;
    }
}

__global__ void k_zero_i(int *p, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_zero_d2(double2 *p, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_double2(0.0, 0.0);
}

// grid set-up of computeQ_GPU_Grid / computeFH_GPU_Grid for an image of
// n[0] x n[1] x n[2] (computeQ.cmem.cu:433-498, computeFH.cmem.cu:495-543):
// gridOS rounded in float so x/y grids are even (2D) or z a multiple of 4
// (3D), Beatty's beta, G = ceil(os n) with x and y bumped to even
GridGeom grid_geometry(const int n[3], float gridOS) {
    GridGeom g{};
    for (int a = 0; a < 3; ++a) g.n[a] = n[a];
    g.d3 = n[2] > 1;
    float os = gridOS;
    if (!g.d3)
        os = 2.0f * std::ceil((os * (float)n[0]) / 2.0f) / (float)n[0];
    else
        os = 4.0f * std::ceil((os * (float)n[2]) / 4.0f) / (float)n[2];
    g.os = os;
    g.inv_os = 1.0f / os;
    g.half = KB_WIDTH * os / 2.0f;
    const double o = (double)os, W = KB_WIDTH;
    g.beta = (float)(M_PI * std::sqrt((o - 0.5) * (o - 0.5) * (W * W * 4.0) / (o * o) - 0.8));
    g.G[0] = (int)std::ceil(os * (float)n[0]);
    g.G[1] = (int)std::ceil(os * (float)n[1]);
    g.G[0] += g.G[0] % 2;
    g.G[1] += g.G[1] % 2;
    g.G[2] = g.d3 ? (int)std::ceil(os * (float)n[2]) : 1;
    return g;
}

// pointers carved from one workspace slot
struct Carve {
    char *p;
    template <class T>
    T *take(size_t count) {
        T *r = (T *)p;
        p += (count * sizeof(T) + 255) & ~(size_t)255;
        return r;
    }
};

// Sorted samples of one trajectory on one grid: cell ranges, positions and
// segment weights.
struct Binned {
    int nbuck;        // time buckets (segments + 1)
    int *start;       // [nbuck * ncells + 1]
    int *perm;        // [m] sample index of sorted position
    float4 *pos;      // [m] sorted
    float2 *w;        // [m] sorted segment weights
};

// the radix sort's histogram of m elements, and the block sums of an
// exclusive scan of n ints
long rs_blocks(int m) { return (m + RS_BLOCK - 1) / RS_BLOCK; }
long scan_blocks(long n) { return (n + GT * 4 - 1) / (GT * 4); }

// keys: time bucket x grid cell; nseg + 1 buckets
long key_range(long ncells, int nseg) { return ncells * (nseg + 1); }

size_t binned_bytes(int m, long ncells, int nseg) {
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const long nk = key_range(ncells, nseg);
    const long nh = (long)RS_BINS * rs_blocks(m);
    const long nb = std::max(scan_blocks(nk), scan_blocks(nh));
    return al((nk + 1) * 4) * 2 + al(nb * 4) + al(16) + al((size_t)m * 4) * 4 +
           al((size_t)(nh + 1) * 4) * 2 + al((size_t)m * 16) * 2 + al((size_t)m * 8);
}

// exclusive scan of n ints in three launches (out has n + 1 entries: the
// total last); bsum holds scan_blocks(n) ints
void scan_ints(hipStream_t st, const int *in, int *out, long n, int *bsum, int *total) {
    const long nb = scan_blocks(n);
    hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(GT), 0, st, in, (int)n, out, bsum);
    hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(GT), 0, st, bsum, (int)nb, total);
    hipLaunchKernelGGL(k_scan_add, dim3(ceil_div(n + 1, 256)), dim3(256), 0, st, out, (int)n, bsum,
                       total);
}

int bin_samples(imp_ctx *ctx, Carve &cv, const GridGeom &g, int m, const float *kx,
                const float *ky, const float *kz, float kscale, const float *t, int nseg,
                float tau, float t0, Binned &B) {
    const hipStream_t st = ctx->stream;
    const long ncells = (long)g.G[0] * g.G[1] * g.G[2];
    const long nk = key_range(ncells, nseg);
    if (nk >= (1L << 30))
        return fail(IMP_EUNSUPPORTED, "gridding: %ld time buckets x grid cells (at most 2^30)", nk);
    const long nrb = rs_blocks(m), nh = (long)RS_BINS * nrb;
    const long nb = std::max(scan_blocks(nk), scan_blocks(nh));
    int *count = cv.take<int>(nk + 1);
    B.start = cv.take<int>(nk + 1);
    int *bsum = cv.take<int>(nb);
    int *total = cv.take<int>(4);
    int *key[2] = {cv.take<int>(m), cv.take<int>(m)};
    int *val[2] = {cv.take<int>(m), cv.take<int>(m)};
    int *hist = cv.take<int>(nh + 1);
    int *offs = cv.take<int>(nh + 1);
    float4 *pos_raw = cv.take<float4>(m);
    B.pos = cv.take<float4>(m);
    B.w = cv.take<float2>(m);
    hipLaunchKernelGGL(k_zero_i, dim3(ceil_div(nk + 1, 256)), dim3(256), 0, st, count, nk + 1);
    hipLaunchKernelGGL(k_grid_cells, dim3(ceil_div(m, 256)), dim3(256), 0, st, m, kx, ky, kz, t,
                       kscale, g, t0, tau, nseg, pos_raw, key[0], count);
    scan_ints(st, count, B.start, nk, bsum, total);
    hipLaunchKernelGGL(k_iota, dim3(ceil_div(m, 256)), dim3(256), 0, st, m, val[0]);
    IMP_CHECK_LAUNCH();
    // keys are < nk: sort on the bits that can be set
    B.nbuck = nseg + 1;
    int bits = 1;
    while (bits < 31 && (1L << bits) < nk) ++bits;
    int cur = 0;
    for (int shift = 0; shift < bits; shift += 4, cur ^= 1) {
        hipLaunchKernelGGL(k_rs_count, dim3((unsigned)nrb), dim3(RS_BLOCK), 0, st, key[cur], m,
                           shift, (int)nrb, hist);
        scan_ints(st, hist, offs, nh, bsum, total);
        hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)nrb), dim3(RS_BLOCK), 0, st, key[cur],
                           val[cur], m, shift, (int)nrb, offs, key[cur ^ 1], val[cur ^ 1]);
        IMP_CHECK_LAUNCH();
    }
    B.perm = val[cur];
    hipLaunchKernelGGL(k_grid_samples, dim3(ceil_div(m, 256)), dim3(256), 0, st, m, B.perm,
                       pos_raw, t, nseg, tau, t0, B.pos, B.w);
    IMP_CHECK_LAUNCH();
    return IMP_OK;
}

// line-transform schedule of one axis
struct AxisPlan {
    int G, R, P, logP, n, logLPW;
    float2 *tw;
    float *dp;
};

size_t axis_bytes(int G, int n) {
    return (((size_t)G * 8 + 255) & ~(size_t)255) + (((size_t)n * 4 + 255) & ~(size_t)255);
}

int plan_axis(imp_ctx *ctx, Carve &cv, const GridGeom &g, int a, AxisPlan &ap) {
    ap.G = g.G[a];
    ap.n = g.n[a];
    int P = 1, logP = 0;
    while (ap.G % (2 * P) == 0) {
        P *= 2;
        ++logP;
    }
    ap.P = P;
    ap.logP = logP;
    ap.R = ap.G / P;
    // lines per workgroup: a power of two, at most 16 and 64 KB of LDS
    ap.logLPW = 0;
    while (ap.logLPW < 4 && (2L << ap.logLPW) * ap.G * (long)sizeof(float2) <= 65536) ++ap.logLPW;
    ap.tw = cv.take<float2>(ap.G);
    ap.dp = cv.take<float>(ap.n);
    hipLaunchKernelGGL(k_twiddle_g, dim3(ceil_div(ap.G, 256)), dim3(256), 0, ctx->stream, ap.G,
                       ap.tw);
    hipLaunchKernelGGL(k_deapod, dim3(ceil_div(ap.n, 256)), dim3(256), 0, ctx->stream, ap.n,
                       (double)g.os, (double)g.beta, ap.dp);
    IMP_CHECK_LAUNCH();
    return IMP_OK;
}

int run_axis(imp_ctx *ctx, const AxisPlan &ap, const float2 *in, float2 *out, long nlines,
             long A) {
    if (nlines > INT32_MAX / 2 || A > INT32_MAX / 2 || (long)ap.n * A > INT32_MAX)
        return fail(IMP_EUNSUPPORTED, "gridding: %ld lines of %ld too many", nlines, A);
    const int lpw = 1 << ap.logLPW;
    const unsigned wgs = (unsigned)((nlines + lpw - 1) / lpw);
    const size_t lds = (size_t)lpw * ap.G * sizeof(float2);
    hipLaunchKernelGGL(k_line_dft, dim3(wgs), dim3(GT), lds, ctx->stream, in, out, ap.G, ap.R,
                       ap.P, ap.logP, ap.n, (int)A, (int)nlines, ap.logLPW,
                       1.0f / (float)ap.R, ap.tw, ap.dp);
    IMP_CHECK_LAUNCH();
    return IMP_OK;
}

// every axis of `planes` grids [gy][gx][gz] -> images [ny][nx][nz]; the
// result ends in *res (one of a / b)
int transform_planes(imp_ctx *ctx, const GridGeom &g, AxisPlan ax[3], int planes, float2 *a,
                     float2 *b, float2 **res) {
    // current layout dims, outermost first; the contiguous axis is last
    const int nd = g.d3 ? 3 : 2;
    // axis order of the passes: contiguous axis first (z in 3D, x in 2D)
    const int order3[3] = {2, 0, 1}, order2[2] = {0, 1};
    const int *order = g.d3 ? order3 : order2;
    // sizes of the current layout (all axes), starting from the grid
    int cur[3] = {g.G[0], g.G[1], g.G[2]};
    float2 *src = a, *dst = b;
    for (int p = 0; p < nd; ++p) {
        const int axis = order[p];
        long per_plane = 1;
        for (int q = 0; q < 3; ++q) per_plane *= cur[q];
        const long A = per_plane / cur[axis];
        IMP_TRY(run_axis(ctx, ax[axis], src, dst, (long)planes * A, A));
        cur[axis] = g.n[axis];
        std::swap(src, dst);
    }
    *res = src;
    return IMP_OK;
}

int check_traj(const GridGeom &g) {
    for (int a = 0; a < (g.d3 ? 3 : 2); ++a)
        if (g.G[a] > 4096 || g.G[a] < 2)
            return fail(IMP_EUNSUPPORTED, "gridding: grid of %d points along axis %d (2..4096)",
                        g.G[a], a);
    return IMP_OK;
}

void host_t_ends(imp_ctx *ctx, const float *t, int m, float *tt, int *st) {
    *st = IMP_OK;
    if (hipMemcpyAsync(&tt[0], t, sizeof(float), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(&tt[1], t + (m - 1), sizeof(float), hipMemcpyDeviceToHost, ctx->stream) !=
            hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        *st = fail(IMP_EHIP, "gridding: %s", hipGetErrorString(hipGetLastError()));
}

// gather of segment l: light points one per lane, then the listed heavy
// points one per wave (heavy holds ncells ints, nheavy one)
template <int C>
void launch_gather(imp_ctx *ctx, const GridGeom &g, const Binned &B, const float2 *data, int l,
                   float2 *grid, size_t plane, int *heavy, int *nheavy) {
    const long total = (long)g.G[0] * g.G[1] * g.G[2];
    hipLaunchKernelGGL(k_zero_i, dim3(1), dim3(256), 0, ctx->stream, nheavy, 1L);
    hipLaunchKernelGGL(k_grid_gather<C>, dim3((unsigned)ceil_div(total, GT)), dim3(GT), 0,
                       ctx->stream, g, B.start, B.nbuck, B.pos, B.w, data, l, grid, plane, heavy,
                       nheavy);
    hipLaunchKernelGGL(k_grid_gather_heavy<C>, dim3(4 * (unsigned)ctx->num_cus), dim3(GT), 0,
                       ctx->stream, g, B.start, B.nbuck, B.pos, B.w, data, l, grid, plane, heavy,
                       nheavy);
}

template <int C>
void launch_data(imp_ctx *ctx, int m, int ncoils, const int *perm, const float *d_r,
                 const float *d_i, float2 *out) {
    hipLaunchKernelGGL(k_grid_data<C>, dim3(ceil_div(m, 256)), dim3(256), 0, ctx->stream, m,
                       ncoils, perm, d_r, d_i, out);
}

}  // namespace

int gridding_q(imp_ctx *ctx, int num_k, const float *kx, const float *ky, const float *kz,
               const float *t, int Nx, int Ny, int Nz, int ntime_segs, float gridOS, float *q_r,
               float *q_i) {
    if (!ctx || num_k <= 0 || !kx || !ky || !t || !q_r || !q_i || Nx <= 0 || Ny <= 0 ||
        Nz <= 0 || ntime_segs < 0 || !(gridOS >= 1.0f) || (Nz > 1 && !kz))
        return fail(IMP_EINVAL, "imp_toeplitz_q_gridding: bad arguments");
    const int nseg = ntime_segs + 1;
    const int n[3] = {2 * Nx, 2 * Ny, Nz > 1 ? 2 * Nz : 1};
    const GridGeom g = grid_geometry(n, gridOS);
    IMP_TRY(check_traj(g));
    float tt[2];
    int st = IMP_OK;
    host_t_ends(ctx, t, num_k, tt, &st);
    IMP_TRY(st);
    const float tau = ntime_segs > 0 ? (tt[1] - tt[0]) / (float)ntime_segs : (tt[1] - tt[0]);
    const long ncells = (long)g.G[0] * g.G[1] * g.G[2];
    const long nq = (long)n[0] * n[1] * n[2];
    const size_t gbytes = (size_t)nseg * ncells * sizeof(float2);
    size_t bytes = binned_bytes(num_k, ncells, nseg) + 2 * (((gbytes + 255) & ~(size_t)255)) + 4096 +
                   (((size_t)ncells * 4 + 255) & ~(size_t)255) + 256 +
                   axis_bytes(g.G[0], n[0]) + axis_bytes(g.G[1], n[1]) + axis_bytes(g.G[2], n[2]);
    void *ws = nullptr;
    IMP_TRY(workspace(ctx, WS_GRID, bytes, &ws));
    Carve cv{(char *)ws};
    Binned B{};
    // Q samples sit at 2k on the 2N grid, with unit data (computeQ.cmem.cu:515-525)
    IMP_TRY(bin_samples(ctx, cv, g, num_k, kx, ky, kz, 2.0f, t, nseg, tau, tt[0], B));
    float2 *bufA = cv.take<float2>((size_t)nseg * ncells);
    float2 *bufB = cv.take<float2>((size_t)nseg * ncells);
    int *heavy = cv.take<int>(ncells);
    int *nheavy = cv.take<int>(1);
    AxisPlan ax[3];
    for (int a = 0; a < (g.d3 ? 3 : 2); ++a) IMP_TRY(plan_axis(ctx, cv, g, a, ax[a]));
    const int rec = prof_begin(ctx, "k_grid_gather");
    for (int l = 0; l < nseg; ++l)
        launch_gather<0>(ctx, g, B, nullptr, l, bufA + (size_t)l * ncells, 0, heavy, nheavy);
    prof_end(ctx, rec);
    IMP_CHECK_LAUNCH();
    float2 *res = nullptr;
    const int rec2 = prof_begin(ctx, "k_line_dft");
    IMP_TRY(transform_planes(ctx, g, ax, nseg, bufA, bufB, &res));
    prof_end(ctx, rec2);
    hipLaunchKernelGGL(k_grid_q_out, dim3(ceil_div((long)nseg * nq, 256)), dim3(256), 0,
                       ctx->stream, (long)nseg * nq, res, q_r, q_i);
    IMP_CHECK_LAUNCH();
    IMP_TRY(workspace_release(ctx, WS_GRID, 1ull << 30));
    return IMP_OK;
}

int gridding_fhd(imp_ctx *ctx, int num_k, int ncoils, const float *kx, const float *ky,
                 const float *kz, const float *t, int Nx, int Ny, int Nz, const float *fm,
                 const float *d_r, const float *d_i, const float *s_r, const float *s_i,
                 int ntime_segs, float gridOS, float *out_r, float *out_i) {
    if (!ctx || num_k <= 0 || ncoils <= 0 || !kx || !ky || !t || !d_r || !d_i || !out_r ||
        !out_i || Nx <= 0 || Ny <= 0 || Nz <= 0 || ntime_segs < 0 || !(gridOS >= 1.0f) ||
        (Nz > 1 && !kz) || (s_r == nullptr) != (s_i == nullptr))
        return fail(IMP_EINVAL, "imp_fhd_gridding: bad arguments");
    const int nseg = ntime_segs + 1;
    const int n[3] = {Nx, Ny, Nz};
    const GridGeom g = grid_geometry(n, gridOS);
    IMP_TRY(check_traj(g));
    float tt[2];
    int st = IMP_OK;
    host_t_ends(ctx, t, num_k, tt, &st);
    IMP_TRY(st);
    const float tau = ntime_segs > 0 ? (tt[1] - tt[0]) / (float)ntime_segs : (tt[1] - tt[0]);
    const long ncells = (long)g.G[0] * g.G[1] * g.G[2];
    const long nv = (long)Nx * Ny * Nz;
    const int CP = ncoils <= 1 ? 1 : ncoils <= 2 ? 2 : ncoils <= 4 ? 4 : ncoils <= 8 ? 8
                 : ncoils <= 16 ? 16 : 32;
    const int groups = ceil_div(ncoils, CP);
    const size_t gbytes = (size_t)CP * ncells * sizeof(float2);
    size_t bytes = binned_bytes(num_k, ncells, nseg) + 2 * ((gbytes + 255) & ~(size_t)255) +
                   (((size_t)ncells * 4 + 255) & ~(size_t)255) + 256 +
                   (((size_t)num_k * CP * 8 + 255) & ~(size_t)255) + nv * 16 + 8192 +
                   axis_bytes(g.G[0], n[0]) + axis_bytes(g.G[1], n[1]) + axis_bytes(g.G[2], n[2]);
    void *ws = nullptr;
    IMP_TRY(workspace(ctx, WS_GRID, bytes, &ws));
    Carve cv{(char *)ws};
    Binned B{};
    IMP_TRY(bin_samples(ctx, cv, g, num_k, kx, ky, kz, 1.0f, t, nseg, tau, tt[0], B));
    float2 *bufA = cv.take<float2>((size_t)CP * ncells);
    float2 *bufB = cv.take<float2>((size_t)CP * ncells);
    int *heavy = cv.take<int>(ncells);
    int *nheavy = cv.take<int>(1);
    float2 *data = cv.take<float2>((size_t)num_k * CP);
    double2 *acc = cv.take<double2>(nv);
    AxisPlan ax[3];
    for (int a = 0; a < (g.d3 ? 3 : 2); ++a) IMP_TRY(plan_axis(ctx, cv, g, a, ax[a]));
    hipLaunchKernelGGL(k_zero_d2, dim3(ceil_div(nv, 256)), dim3(256), 0, ctx->stream, acc, nv);
    for (int grp = 0; grp < groups; ++grp) {
        const int c0 = grp * CP, cn = std::min(CP, ncoils - c0);
        const float *dr = d_r + (size_t)c0 * num_k, *di = d_i + (size_t)c0 * num_k;
        switch (CP) {
            case 1: launch_data<1>(ctx, num_k, cn, B.perm, dr, di, data); break;
            case 2: launch_data<2>(ctx, num_k, cn, B.perm, dr, di, data); break;
            case 4: launch_data<4>(ctx, num_k, cn, B.perm, dr, di, data); break;
            case 8: launch_data<8>(ctx, num_k, cn, B.perm, dr, di, data); break;
            case 16: launch_data<16>(ctx, num_k, cn, B.perm, dr, di, data); break;
            default: launch_data<32>(ctx, num_k, cn, B.perm, dr, di, data); break;
        }
        IMP_CHECK_LAUNCH();
        for (int l = 0; l < nseg; ++l) {
            const int rec = prof_begin(ctx, "k_grid_gather");
            switch (CP) {
                case 1: launch_gather<1>(ctx, g, B, data, l, bufA, ncells, heavy, nheavy); break;
                case 2: launch_gather<2>(ctx, g, B, data, l, bufA, ncells, heavy, nheavy); break;
                case 4: launch_gather<4>(ctx, g, B, data, l, bufA, ncells, heavy, nheavy); break;
                case 8: launch_gather<8>(ctx, g, B, data, l, bufA, ncells, heavy, nheavy); break;
                case 16: launch_gather<16>(ctx, g, B, data, l, bufA, ncells, heavy, nheavy); break;
                default: launch_gather<32>(ctx, g, B, data, l, bufA, ncells, heavy, nheavy); break;
            }
            prof_end(ctx, rec);
            IMP_CHECK_LAUNCH();
            float2 *res = nullptr;
            const int rec2 = prof_begin(ctx, "k_line_dft");
            IMP_TRY(transform_planes(ctx, g, ax, cn, bufA, bufB, &res));
            prof_end(ctx, rec2);
            // expfm argument fm (tau l + t0) in float (recon.cu:1281-1285)
            const float tl = tau * (float)l + tt[0];
            hipLaunchKernelGGL(k_grid_fhd_accum, dim3(ceil_div(nv, 256)), dim3(256), 0,
                               ctx->stream, nv, cn, res, (size_t)nv,
                               s_r ? s_r + (size_t)c0 * nv : nullptr,
                               s_i ? s_i + (size_t)c0 * nv : nullptr, fm, tl, acc);
            IMP_CHECK_LAUNCH();
        }
    }
    hipLaunchKernelGGL(k_grid_fhd_out, dim3(ceil_div(nv, 256)), dim3(256), 0, ctx->stream, nv,
                       acc, out_r, out_i);
    IMP_CHECK_LAUNCH();
    IMP_TRY(workspace_release(ctx, WS_GRID, 1ull << 30));
    return IMP_OK;
}

}  // namespace imp

extern "C" int imp_toeplitz_q_gridding(imp_ctx *ctx, int num_k, const float *kx, const float *ky,
                                       const float *kz, const float *t, int Nx, int Ny, int Nz,
                                       int ntime_segs, float gridOS, float *q_r, float *q_i) {
    return imp::gridding_q(ctx, num_k, kx, ky, kz, t, Nx, Ny, Nz, ntime_segs, gridOS, q_r, q_i);
}

extern "C" int imp_fhd_gridding(imp_ctx *ctx, int num_k, int ncoils, const float *kx,
                                const float *ky, const float *kz, const float *t, int Nx, int Ny,
                                int Nz, const float *fm, const float *d_r, const float *d_i,
                                const float *s_r, const float *s_i, int ntime_segs, float gridOS,
                                float *out_r, float *out_i) {
    return imp::gridding_fhd(ctx, num_k, ncoils, kx, ky, kz, t, Nx, Ny, Nz, fm, d_r, d_i, s_r,
                             s_i, ntime_segs, gridOS, out_r, out_i);
}
