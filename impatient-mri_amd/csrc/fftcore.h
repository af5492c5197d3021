// Register-staged Stockham FFT building blocks for gfx950.
//
// A length-L line is transformed by T = L / E threads that each hold E = 16
// values (E = L for L < 16).  Stage s uses radix R_s from a fixed plan
// (16 first, then 16 / 8 / 4 / 2); every stage gives each thread
// Q / T = L / (R T) butterflies, so a thread always owns E values and the
// element sets it owns before and after a stage follow from (t, b, r)
// alone.  Between stages the values go through LDS; the first stage may read
// global memory directly and the last stage may write it directly.
//
// Stockham step with span NS (product of the earlier radices), Q = L / R:
//   for j in [0, Q): k = j mod NS
//     v[r] = x[j + r Q] * w^(r k L / (NS R)),  w = e^(-+2 pi i / L)
//     y[(j - k) R + k + r NS] = DFT_R(v)[r]
// After the last stage y is the DFT of x in natural order.
#pragma once

#include <hip/hip_runtime.h>

namespace imp {
namespace fftc {

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
// multiply by e^(-+ i pi / 4) (forward / inverse)
template <bool INV>
__device__ __forceinline__ float2 rot45(float2 a) {
    constexpr float h = 0.70710678118654752f;
    return INV ? make_float2(h * (a.x - a.y), h * (a.x + a.y))
               : make_float2(h * (a.x + a.y), h * (a.y - a.x));
}
template <bool INV>
__device__ __forceinline__ float2 rot135(float2 a) {  // e^(-+ 3 i pi / 4)
    constexpr float h = 0.70710678118654752f;
    return INV ? make_float2(-h * (a.x + a.y), h * (a.x - a.y))
               : make_float2(h * (a.y - a.x), -h * (a.x + a.y));
}

// In-register DFT of size R in natural order: v[k] <- sum_n v[n] w^(nk),
// w = e^(-2 pi i / R) (forward) or e^(+2 pi i / R) (inverse).
template <int R, bool INV>
__device__ __forceinline__ void dft(float2 *v);

template <>
__device__ __forceinline__ void dft<1, false>(float2 *) {}
template <>
__device__ __forceinline__ void dft<1, true>(float2 *) {}

template <bool INV>
__device__ __forceinline__ void dft2(float2 *v) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}
template <>
__device__ __forceinline__ void dft<2, false>(float2 *v) { dft2<false>(v); }
template <>
__device__ __forceinline__ void dft<2, true>(float2 *v) { dft2<true>(v); }

template <bool INV>
__device__ __forceinline__ void dft4(float2 *v) {
    const float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    const float2 s13 = cadd(v[1], v[3]), d13 = rot90<INV>(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
}
template <>
__device__ __forceinline__ void dft<4, false>(float2 *v) { dft4<false>(v); }
template <>
__device__ __forceinline__ void dft<4, true>(float2 *v) { dft4<true>(v); }

// radix-2 decimation in time: X[k] = E[k] + w^k O[k], X[k + R/2] = E[k] - w^k O[k]
template <bool INV>
__device__ __forceinline__ void dft8(float2 *v) {
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    dft4<INV>(e);
    dft4<INV>(o);
    o[1] = rot45<INV>(o[1]);
    o[2] = rot90<INV>(o[2]);
    o[3] = rot135<INV>(o[3]);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = cadd(e[k], o[k]);
        v[k + 4] = csub(e[k], o[k]);
    }
}
template <>
__device__ __forceinline__ void dft<8, false>(float2 *v) { dft8<false>(v); }
template <>
__device__ __forceinline__ void dft<8, true>(float2 *v) { dft8<true>(v); }

template <bool INV>
__device__ __forceinline__ void dft16(float2 *v) {
    // 16 = 4 x 4: DFT-4 over n = a + 4 c (c), twiddle w16^(a k1), DFT-4 over a
    float2 x[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
#pragma unroll
        for (int c = 0; c < 4; ++c) x[a][c] = v[a + 4 * c];
        dft4<INV>(x[a]);  // x[a][k1]
    }
    // twiddles w16^(a k1) for a, k1 in 1..3 (exponents 1, 2, 3, 4, 6, 9);
    // w16^m = cos(pi m / 8) -+ i sin(pi m / 8)
    constexpr float c1 = 0.92387953251128676f, s1 = 0.38268343236508977f;
    constexpr float sg = INV ? 1.f : -1.f;
    const float2 w1 = make_float2(c1, sg * s1), w3 = make_float2(s1, sg * c1),
                 w9 = make_float2(-c1, -sg * s1);
    x[1][1] = cmul(x[1][1], w1);
    x[1][2] = rot45<INV>(x[1][2]);
    x[1][3] = cmul(x[1][3], w3);
    x[2][1] = rot45<INV>(x[2][1]);
    x[2][2] = rot90<INV>(x[2][2]);
    x[2][3] = rot135<INV>(x[2][3]);
    x[3][1] = cmul(x[3][1], w3);
    x[3][2] = rot135<INV>(x[3][2]);
    x[3][3] = cmul(x[3][3], w9);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        float2 y[4] = {x[0][k1], x[1][k1], x[2][k1], x[3][k1]};
        dft4<INV>(y);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = y[k2];
    }
}
template <>
__device__ __forceinline__ void dft<16, false>(float2 *v) { dft16<false>(v); }
template <>
__device__ __forceinline__ void dft<16, true>(float2 *v) { dft16<true>(v); }

template <bool INV>
__device__ __forceinline__ void dft3(float2 *v) {
    constexpr float r3 = 0.86602540378443865f;  // sqrt(3) / 2
    const float2 s = cadd(v[1], v[2]);
    const float2 d = rot90<INV>(csub(v[1], v[2]));
    const float2 t = make_float2(v[0].x - 0.5f * s.x, v[0].y - 0.5f * s.y);
    v[0] = cadd(v[0], s);
    v[1] = make_float2(t.x + r3 * d.x, t.y + r3 * d.y);
    v[2] = make_float2(t.x - r3 * d.x, t.y - r3 * d.y);
}
template <>
__device__ __forceinline__ void dft<3, false>(float2 *v) { dft3<false>(v); }
template <>
__device__ __forceinline__ void dft<3, true>(float2 *v) { dft3<true>(v); }

template <bool INV>
__device__ __forceinline__ void dft12(float2 *v) {
    // 12 = 4 x 3: DFT-3 over n = a + 4 c (c), twiddle w12^(a k1), DFT-4 over a;
    // output k = k1 + 3 k2
    float2 x[4][3];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
#pragma unroll
        for (int c = 0; c < 3; ++c) x[a][c] = v[a + 4 * c];
        dft3<INV>(x[a]);
    }
    // w12^m = cos(pi m / 6) -+ i sin(pi m / 6), m = a k1 in {1, 2, 2, 4, 3, 6}
    constexpr float r3 = 0.86602540378443865f;
    constexpr float sg = INV ? 1.f : -1.f;
    const float2 w1 = make_float2(r3, sg * 0.5f), w2 = make_float2(0.5f, sg * r3),
                 w4 = make_float2(-0.5f, sg * r3);
    x[1][1] = cmul(x[1][1], w1);
    x[1][2] = cmul(x[1][2], w2);
    x[2][1] = cmul(x[2][1], w2);
    x[2][2] = cmul(x[2][2], w4);
    x[3][1] = rot90<INV>(x[3][1]);
    x[3][2] = make_float2(-x[3][2].x, -x[3][2].y);
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1) {
        float2 y[4] = {x[0][k1], x[1][k1], x[2][k1], x[3][k1]};
        dft4<INV>(y);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[k1 + 3 * k2] = y[k2];
    }
}
template <>
__device__ __forceinline__ void dft<12, false>(float2 *v) { dft12<false>(v); }
template <>
__device__ __forceinline__ void dft<12, true>(float2 *v) { dft12<true>(v); }

// ------------------------------------------------------------------ plans
// Supported lengths: 2^k (2..4096) and 3 * 2^k (12..3072).
// Powers of two: E = 16 values per thread (E = L below 16), radix 16 while
// at least 4 bits remain, then the rest (2, 4, 8 or 16).
// Lengths 3 * 2^k: E = 12, a radix-12 stage first, then radix 4, and a
// final radix 2 when an odd power of two remains; every radix divides 12,
// so each stage gives each of the T = L / 12 threads whole butterflies.
__host__ __device__ constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }
__host__ __device__ constexpr bool has3(int L) { return L % 3 == 0; }
__host__ __device__ constexpr int vals_per_thread(int L) {
    return has3(L) ? (L < 12 ? L : 12) : (L < 16 ? L : 16);
}
__host__ __device__ constexpr int threads_per_line(int L) { return L / vals_per_thread(L); }
// Threads per workgroup of the row-transform passes: 512 (8 lines of 1024
// points, whose LDS twiddle table then still leaves two workgroups per CU)
// from 512-point lines up, 256 below.
__host__ __device__ constexpr int row_threads(int L) { return L >= 512 ? 512 : 256; }
__host__ __device__ constexpr int num_stages(int L) {
    return has3(L) ? (L <= 12 ? 1 : 1 + (ilog2c(L / 12) + 1) / 2)
                   : (L <= 16 ? 1 : (ilog2c(L) + 3) / 4);
}
__host__ __device__ constexpr int stage_radix(int L, int s) {
    return has3(L) ? (L <= 12 ? L : (s == 0 ? 12 : (ilog2c(L / 12) - 2 * (s - 1) >= 2 ? 4 : 2)))
                   : (L <= 16 ? L : (s + 1 < num_stages(L) ? 16 : (1 << (ilog2c(L) - 4 * s))));
}
__host__ __device__ constexpr int stage_span(int L, int s) {  // NS: product of earlier radices
    return s == 0 ? 1 : stage_span(L, s - 1) * stage_radix(L, s - 1);
}
__host__ __device__ constexpr bool fft_length_supported(int L) {
    return L >= 2 && (has3(L) ? (L >= 12 && L <= 3072 && ((L / 3) & (L / 3 - 1)) == 0)
                              : (L <= 4096 && (L & (L - 1)) == 0));
}

// LDS index of element i of a line: one pad word every 16 elements keeps the
// strided stage writes within two-way bank conflicts.
__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }
__host__ __device__ constexpr int line_lds(int L) { return L + L / 16; }

// One Stockham stage for the thread with line-local index t.  LD(b, r, i)
// returns input element i (= j + r Q, j = t + T b); ST(b, r, i, v) stores
// output element i; b and r are compile-time after unrolling.  Twiddles come
// from tw[m] = exp(-2 pi i m / L).
//   FUSED = false: every load and butterfly of the thread, then MID() (a
//                  barrier when the stage reads and writes the same LDS
//                  lines), then every store.
//   FUSED = true:  MID() first, then load / butterfly / store one butterfly
//                  at a time (loads and stores must not alias); keeps fewer
//                  values live.
template <int L, int S, bool INV, bool FUSED, class LD, class MID, class ST>
__device__ __forceinline__ void stage(int t, const float2 *__restrict__ tw, LD ld, MID mid,
                                      ST st) {
    constexpr int R = stage_radix(L, S);
    constexpr int NS = stage_span(L, S);
    constexpr int T = threads_per_line(L);
    constexpr int Q = L / R;
    constexpr int NB = Q / T;
    static_assert(NB >= 1 && Q % T == 0, "radix plan does not fit the thread layout");
    auto butterfly = [&](int b, float2 *v) {
        const int j = t + T * b;
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = ld(b, r, j + r * Q);
        if constexpr (NS > 1) {
            const int k = j % NS;
            const int step = (L / (NS * R)) * k;
#pragma unroll
            for (int r = 1; r < R; ++r) {
                float2 w = tw[r * step];
                if (INV) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        dft<R, INV>(v);
    };
    auto store = [&](int b, const float2 *v) {
        const int j = t + T * b;
        const int k = NS > 1 ? j % NS : 0;
#pragma unroll
        for (int r = 0; r < R; ++r) st(b, r, (j - k) * R + k + r * NS, v[r]);
    };
    if constexpr (FUSED) {
        mid();
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            float2 v[R];
            butterfly(b, v);
            store(b, v);
        }
    } else {
        float2 v[NB][R];
#pragma unroll
        for (int b = 0; b < NB; ++b) butterfly(b, v[b]);
        mid();
#pragma unroll
        for (int b = 0; b < NB; ++b) store(b, v[b]);
    }
}

// Register index m of output element i after the last stage (i = t + T m).
template <int L>
__host__ __device__ constexpr int last_reg(int b, int r) {
    return b + (L / (stage_radix(L, num_stages(L) - 1) * threads_per_line(L))) * r;
}
// Register index m of input element i of the first stage (i = t + T m).
template <int L>
__host__ __device__ constexpr int first_reg(int b, int r) {
    return b + (L / (stage_radix(L, 0) * threads_per_line(L))) * r;
}

}  // namespace fftc
}  // namespace imp
