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
//   - samples are sorted by grid cell (a counting sort made deterministic by
//     ordering each cell's samples by index), so each grid point gathers
//     from contiguous sample ranges with no atomics on the grid;
//   - one gather launch grids every coil of a segment at once (each sample's
//     kernel weight serves all coils);
//   - the DFT of step 2 runs per axis on lines of length G = R P (P the
//     largest power of two dividing G): R radix-2 FFTs of length P in LDS,
//     then the R-term combination evaluated only at the n kept outputs, with
//     step 3 folded in; every pass writes its output with the transformed
//     axis outermost, so after one pass per axis the layout is the image's.
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

// shifted positions s = os (k + n/2) (float, as the reference computes them)
// and the grid cell ((cy Gx) + cx) Gz + cz of every sample
__global__ void k_grid_cells(int m, const float *kx, const float *ky, const float *kz,
                             float kscale, GridGeom g, float4 *pos, int *cell, int *count) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const float sx = g.os * (kscale * kx[j] + (float)g.n[0] / 2.0f);
    const float sy = g.os * (kscale * ky[j] + (float)g.n[1] / 2.0f);
    const float sz = g.d3 ? g.os * (kscale * kz[j] + (float)g.n[2] / 2.0f) : 0.f;
    const int cx = min(g.G[0] - 1, max(0, (int)floorf(sx)));
    const int cy = min(g.G[1] - 1, max(0, (int)floorf(sy)));
    const int cz = g.d3 ? min(g.G[2] - 1, max(0, (int)floorf(sz))) : 0;
    const int c = ((cy * g.G[0]) + cx) * g.G[2] + cz;
    pos[j] = make_float4(sx, sy, sz, 0.f);
    cell[j] = c;
    atomicAdd(&count[c], 1);
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

// scatter sample j to its cell's range (any order within a cell)
__global__ void k_cell_scatter(int m, const int *cell, int *cursor, int *perm) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < m) perm[atomicAdd(&cursor[cell[j]], 1)] = j;
}

// order each cell's samples by sample index (cells hold a handful of
// samples), so the gather sums in a fixed order on every run
__global__ void k_cell_sort(int ncells, const int *start, int *perm) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    const int b = start[c], e = start[c + 1];
    for (int i = b + 1; i < e; ++i) {
        const int v = perm[i];
        int k = i - 1;
        while (k >= b && perm[k] > v) {
            perm[k + 1] = perm[k];
            --k;
        }
        perm[k + 1] = v;
    }
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

// One thread per grid point: sum over the samples within the kernel window
// of segment l (weight a_l), for C coils (C = 0: unit data, one plane).
// Grid layout per coil: [gy][gx][gz], written at grid + c * plane.
template <int C>
__global__ __launch_bounds__(GT) void k_grid_gather(GridGeom g, const int *start,
                                                    const float4 *pos, const float2 *w,
                                                    const float2 *data, int l, float2 *grid,
                                                    size_t plane) {
    constexpr int NC = C > 0 ? C : 1;
    const long id = (long)blockIdx.x * GT + threadIdx.x;
    const long total = (long)g.G[0] * g.G[1] * g.G[2];
    if (id >= total) return;
    // fastest index along the innermost cell axis (z in 3D, x in 2D)
    int gx, gy, gz;
    if (g.d3) {
        gz = (int)(id % g.G[2]);
        gx = (int)((id / g.G[2]) % g.G[0]);
        gy = (int)(id / ((long)g.G[2] * g.G[0]));
    } else {
        gz = 0;
        gx = (int)(id % g.G[0]);
        gy = (int)(id / g.G[0]);
    }
    const int r_lo = (int)ceilf(g.half), r_hi = (int)floorf(g.half);
    float acc_r[NC], acc_i[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc_r[c] = acc_i[c] = 0.f;
    const int y0 = max(0, gy - r_lo), y1 = min(g.G[1] - 1, gy + r_hi);
    const int x0 = max(0, gx - r_lo), x1 = min(g.G[0] - 1, gx + r_hi);
    const int z0 = g.d3 ? max(0, gz - r_lo) : 0, z1 = g.d3 ? min(g.G[2] - 1, gz + r_hi) : 0;
    // outer cell axes: (y, x) in 3D with z contiguous; (y) in 2D with x contiguous
    const int o1 = g.d3 ? x0 : 0, o1e = g.d3 ? x1 : 0;
    for (int cy = y0; cy <= y1; ++cy) {
        for (int cx = o1; cx <= o1e; ++cx) {
            int b, e;
            if (g.d3) {
                const long row = ((long)cy * g.G[0] + cx) * g.G[2];
                b = start[row + z0];
                e = start[row + z1 + 1];
            } else {
                const long row = (long)cy * g.G[0];
                b = start[row + x0];
                e = start[row + x1 + 1];
            }
            for (int ks = b; ks < e; ++ks) {
                const float4 p = pos[ks];
                const float2 ww = w[ks];
                const int l0 = __float_as_int(p.w);
                const float a = l == l0 ? ww.x : (l == l0 + 1 ? ww.y : 0.f);
                if (a == 0.f) continue;
                float kw = kb_weight(p.x, gx, g.inv_os, g.beta);
                if (kw == 0.f) continue;
                kw *= kb_weight(p.y, gy, g.inv_os, g.beta);
                if (g.d3) kw *= kb_weight(p.z, gz, g.inv_os, g.beta);
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
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) grid[c * plane + id] = make_float2(acc_r[c], acc_i[c]);
}

// Cropped, shifted inverse DFT along the contiguous axis of every line:
//   out[b][i][a] = dp[i] sum_{g<G} in[b][a][g] e^{+i 2pi g (i - n/2) / G}
// (dp carries (-1)^(i - n/2) of the shift and 1 / (os C(u_i))), for lines
// a < A of batch item b.  G = R P: R radix-2 FFTs of length P in LDS over
// the subsequences g = R q + j, then the R-term combination at each kept i.
// tw[k] = e^{+i 2pi k / G}.  LPW lines per workgroup.
__global__ __launch_bounds__(GT) void k_line_dft(const float2 *__restrict__ in,
                                                 float2 *__restrict__ out, int G, int R, int P,
                                                 int logP, int n, long A, long nlines, int LPW,
                                                 const float2 *__restrict__ tw,
                                                 const float *__restrict__ dp) {
    extern __shared__ float2 lds[];
    const long line0 = (long)blockIdx.x * LPW;
    const int nl = (int)min((long)LPW, nlines - line0);
    // load with bit reversal inside each length-P subsequence
    for (int e = threadIdx.x; e < nl * G; e += GT) {
        const int li = e / G, g = e % G;
        const int j = g % R, q = g / R;
        const int qr = logP ? (int)(__brev((unsigned)q) >> (32 - logP)) : 0;
        lds[li * G + j * P + qr] = in[(line0 + li) * (long)G + g];
    }
    __syncthreads();
    // radix-2 stages: butterfly span h = 1, 2, ..., P/2
    for (int h = 1; h < P; h <<= 1) {
        const int nb = nl * R * (P / 2);
        for (int bi = threadIdx.x; bi < nb; bi += GT) {
            const int k = bi % h;
            const int grp = (bi / h) % (P / (2 * h));
            const int sub = bi / (P / 2);       // (line, j)
            const int base = sub * P + grp * 2 * h + k;
            const float2 wv = tw[(long)k * (G / (2 * h))];
            const float2 u = lds[base], v = lds[base + h];
            const float2 tv = make_float2(v.x * wv.x - v.y * wv.y, v.x * wv.y + v.y * wv.x);
            lds[base] = make_float2(u.x + tv.x, u.y + tv.y);
            lds[base + h] = make_float2(u.x - tv.x, u.y - tv.y);
        }
        __syncthreads();
    }
    // combination at the kept outputs, written with the axis outermost
    for (int e = threadIdx.x; e < nl * n; e += GT) {
        const int li = e % nl, i = e / nl;
        const int nn = i - n / 2;
        const int m = ((nn % P) + P) % P;
        float sr = 0.f, si = 0.f;
        for (int j = 0; j < R; ++j) {
            const float2 y = lds[li * G + j * P + m];
            const long kk = (((long)j * nn) % G + G) % G;
            const float2 wv = tw[kk];
            sr += y.x * wv.x - y.y * wv.y;
            si += y.x * wv.y + y.y * wv.x;
        }
        const long line = line0 + li;
        const long b = line / A, a = line % A;
        const float f = dp[i];
        out[(b * n + i) * A + a] = make_float2(sr * f, si * f);
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
    int *start;       // [ncells + 1]
    int *perm;        // [m] sample index of sorted position
    float4 *pos;      // [m] sorted
    float2 *w;        // [m] sorted segment weights
};

size_t binned_bytes(int m, long ncells) {
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const long nb = (ncells + GT * 4 - 1) / (GT * 4);
    return al((ncells + 1) * 4) * 2 + al(nb * 4) + al(16) + al((size_t)m * 4) * 2 +
           al((size_t)m * 16) * 2 + al((size_t)m * 8);
}

int bin_samples(imp_ctx *ctx, Carve &cv, const GridGeom &g, int m, const float *kx,
                const float *ky, const float *kz, float kscale, const float *t, int nseg,
                float tau, float t0, Binned &B) {
    const hipStream_t st = ctx->stream;
    const long ncells = (long)g.G[0] * g.G[1] * g.G[2];
    const long nb = (ncells + GT * 4 - 1) / (GT * 4);
    int *count = cv.take<int>(ncells + 1);
    B.start = cv.take<int>(ncells + 1);
    int *bsum = cv.take<int>(nb);
    int *total = cv.take<int>(4);
    int *cell = cv.take<int>(m);
    B.perm = cv.take<int>(m);
    float4 *pos_raw = cv.take<float4>(m);
    B.pos = cv.take<float4>(m);
    B.w = cv.take<float2>(m);
    hipLaunchKernelGGL(k_zero_i, dim3(ceil_div(ncells + 1, 256)), dim3(256), 0, st, count,
                       ncells + 1);
    hipLaunchKernelGGL(k_grid_cells, dim3(ceil_div(m, 256)), dim3(256), 0, st, m, kx, ky, kz,
                       kscale, g, pos_raw, cell, count);
    hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(GT), 0, st, count, (int)ncells,
                       B.start, bsum);
    hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(GT), 0, st, bsum, (int)nb, total);
    hipLaunchKernelGGL(k_scan_add, dim3(ceil_div(ncells + 1, 256)), dim3(256), 0, st, B.start,
                       (int)ncells, bsum, total);
    IMP_CHECK_LAUNCH();
    // cursor = copy of start (reuses count)
    IMP_HIP(hipMemcpyAsync(count, B.start, ncells * sizeof(int), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_cell_scatter, dim3(ceil_div(m, 256)), dim3(256), 0, st, m, cell, count,
                       B.perm);
    hipLaunchKernelGGL(k_cell_sort, dim3(ceil_div(ncells, 256)), dim3(256), 0, st, (int)ncells,
                       B.start, B.perm);
    hipLaunchKernelGGL(k_grid_samples, dim3(ceil_div(m, 256)), dim3(256), 0, st, m, B.perm,
                       pos_raw, t, nseg, tau, t0, B.pos, B.w);
    IMP_CHECK_LAUNCH();
    return IMP_OK;
}

// line-transform schedule of one axis
struct AxisPlan {
    int G, R, P, logP, n, LPW;
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
    ap.LPW = std::max(1, std::min(16, 2048 / ap.G));
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
    const unsigned wgs = (unsigned)((nlines + ap.LPW - 1) / ap.LPW);
    const size_t lds = (size_t)ap.LPW * ap.G * sizeof(float2);
    hipLaunchKernelGGL(k_line_dft, dim3(wgs), dim3(GT), lds, ctx->stream, in, out, ap.G, ap.R,
                       ap.P, ap.logP, ap.n, A, nlines, ap.LPW, ap.tw, ap.dp);
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

template <int C>
void launch_gather(imp_ctx *ctx, const GridGeom &g, const Binned &B, const float2 *data, int l,
                   float2 *grid, size_t plane) {
    const long total = (long)g.G[0] * g.G[1] * g.G[2];
    hipLaunchKernelGGL(k_grid_gather<C>, dim3((unsigned)ceil_div(total, GT)), dim3(GT), 0,
                       ctx->stream, g, B.start, B.pos, B.w, data, l, grid, plane);
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
    size_t bytes = binned_bytes(num_k, ncells) + 2 * (((gbytes + 255) & ~(size_t)255)) + 4096 +
                   axis_bytes(g.G[0], n[0]) + axis_bytes(g.G[1], n[1]) + axis_bytes(g.G[2], n[2]);
    void *ws = nullptr;
    IMP_TRY(workspace(ctx, WS_GRID, bytes, &ws));
    Carve cv{(char *)ws};
    Binned B{};
    // Q samples sit at 2k on the 2N grid, with unit data (computeQ.cmem.cu:515-525)
    IMP_TRY(bin_samples(ctx, cv, g, num_k, kx, ky, kz, 2.0f, t, nseg, tau, tt[0], B));
    float2 *bufA = cv.take<float2>((size_t)nseg * ncells);
    float2 *bufB = cv.take<float2>((size_t)nseg * ncells);
    AxisPlan ax[3];
    for (int a = 0; a < (g.d3 ? 3 : 2); ++a) IMP_TRY(plan_axis(ctx, cv, g, a, ax[a]));
    const int rec = prof_begin(ctx, "k_grid_gather");
    for (int l = 0; l < nseg; ++l)
        launch_gather<0>(ctx, g, B, nullptr, l, bufA + (size_t)l * ncells, 0);
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
    size_t bytes = binned_bytes(num_k, ncells) + 2 * ((gbytes + 255) & ~(size_t)255) +
                   (((size_t)num_k * CP * 8 + 255) & ~(size_t)255) + nv * 16 + 8192 +
                   axis_bytes(g.G[0], n[0]) + axis_bytes(g.G[1], n[1]) + axis_bytes(g.G[2], n[2]);
    void *ws = nullptr;
    IMP_TRY(workspace(ctx, WS_GRID, bytes, &ws));
    Carve cv{(char *)ws};
    Binned B{};
    IMP_TRY(bin_samples(ctx, cv, g, num_k, kx, ky, kz, 1.0f, t, nseg, tau, tt[0], B));
    float2 *bufA = cv.take<float2>((size_t)CP * ncells);
    float2 *bufB = cv.take<float2>((size_t)CP * ncells);
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
                case 1: launch_gather<1>(ctx, g, B, data, l, bufA, ncells); break;
                case 2: launch_gather<2>(ctx, g, B, data, l, bufA, ncells); break;
                case 4: launch_gather<4>(ctx, g, B, data, l, bufA, ncells); break;
                case 8: launch_gather<8>(ctx, g, B, data, l, bufA, ncells); break;
                case 16: launch_gather<16>(ctx, g, B, data, l, bufA, ncells); break;
                default: launch_gather<32>(ctx, g, B, data, l, bufA, ncells); break;
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
