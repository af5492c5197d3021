// Direct, field-corrected NUDFT on gfx950: forward F, adjoint F^H (with
// SENSE) and the Toeplitz kernels Q_l.
//
// All three are one computation,
//     out[a][c] = sum_b W[c][b] * exp(i * sign * 2pi * (P_a . B_b)),
// with 4-vectors P_a (owned by a lane) and B_b (streamed), because the phase
// 2pi(k.x) + w(x) t is the dot product of (x, y, z, w/2pi) with (kx, ky, kz, t):
//   adjoint  a = voxel  (x,y,z,fm/2pi), b = sample (kx,ky,kz,t), W = conj(phi) d_c
//   forward  a = sample (kx,ky,kz,t),   b = voxel  (x,y,z,fm/2pi), W = S_c rho
//   Q_l      a = Q-grid point (r,0),    b = sample (kx,ky,kz,0),   W = a_l(t) |phi|^2
// The reference evaluates one coil (or one segment l) per pass and recomputes
// the sincos every time (computeFH.cmem.cu:1166-1178, computeQ.cmem.cu:951-963).
// Here one phase evaluation feeds every right-hand side (coil or segment).
//
// Phase: P . B is formed in FP64 (on gfx950 a non-packed FP32 FMA and an
// FP64 FMA issue at the same rate), reduced with fract to [0, 1) revolutions
// and passed to the hardware sin/cos, which take revolutions (max abs error
// 1.2e-7).  The FP64 phase is exact to ~1e-12 revolutions for |k.x| < 10^3.
//
// Two kernels:
//   k_nudft_valu  few right-hand sides: per-lane complex MACs, streamed
//                 operands read with scalar loads (wave-uniform).
//   k_nudft_mfma  many right-hand sides: the MACs are a complex GEMM
//                 E(points x samples) * W(samples x rhs) on
//                 v_mfma_f32_32x32x16_f16.  Each lane evaluates one phase
//                 per sample for its own point and both cos and sin of it;
//                 v_permlane32_swap turns the two half-waves' values into the
//                 A operands of two 32-point M tiles.  W is pre-packed as B
//                 operands.  FP32 accuracy from fp16 MFMAs: E and W are split
//                 into hi + lo fp16 parts (lo may be subnormal; the MFMA keeps
//                 fp16 subnormals) and E W = Eh Wh + Eh Wl + El Wh (the dropped
//                 El Wl term is below 2^-22 relative).
// Stream positions are stored per step of 8 samples as [x8][y8][w8][z8]
// doubles, so one step's positions arrive in three or four scalar loads.
// Sample sums are blocked and folded into compensated (two-sum) totals;
// chunk partials are reduced in FP64.
#include "common.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <vector>

namespace imp {

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int NU_THREADS = 256;
constexpr int NU_BLK = 32;        // VALU: samples per FP32 block before the compensated fold
constexpr int MF_K = 8;           // MFMA: samples per 32x32x16 step (8 Re + 8 Im rows of K)
constexpr int MF_FOLD = 32;       // MFMA: steps per compensated fold (256 samples)
constexpr int MF_BPAD = 2;        // MFMA: zero steps packed after the last (loads run 2 ahead)
constexpr int MF_COLS = 16;       // MFMA: right-hand sides per 32-column tile (Re + Im)
constexpr int W_EXP = 12;         // W is scaled so max |W| < 2^12
constexpr int MF_WAVE_PTS = 64;   // MFMA: points per wave (two 32-point M tiles)
constexpr int MF_PIPE_VALU = 5;  // MFMA: VALU instructions scheduled after each MFMA

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// doubles per step of MF_K stream positions: x, y, w (+ z in 3D)
__host__ __device__ constexpr int pos_stride(bool d3) { return (d3 ? 4 : 3) * MF_K; }

// Stream positions are read through the constant address space: they are
// never written while a kernel runs, and the compiler then loads each
// step's (wave-uniform) positions into scalar registers even where the
// pointer comes from a job table instead of a restrict-qualified argument.
typedef const __attribute__((address_space(4))) double *cdouble_ptr;

struct alignas(16) BPair {
    half8 hi;
    half8 lo;
};

__device__ __forceinline__ float rev_fract(double u) { return (float)__builtin_amdgcn_fract(u); }

typedef _Float16 half2v __attribute__((ext_vector_type(2)));

// fp16 hi/lo split of 8 FP32 values, packed as 4 dwords each: the hi parts
// of a pair come from one v_cvt_pkrtz (any rounding of hi works) and each
// lo part is v - hi of that same packed hi, read back through op_sel, so
// hi + lo stays within a lo ulp of v.
__device__ __forceinline__ void split8(const float (&v)[8], float one, u32x4 &hi, u32x4 &lo) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const half2v h =
            __builtin_bit_cast(half2v, __builtin_amdgcn_cvt_pkrtz(v[2 * i], v[2 * i + 1]));
        half2v l;
        l[0] = (_Float16)__builtin_fmaf(v[2 * i], one, -(float)h[0]);
        l[1] = (_Float16)__builtin_fmaf(v[2 * i + 1], one, -(float)h[1]);
        hi[i] = __builtin_bit_cast(unsigned int, h);
        lo[i] = __builtin_bit_cast(unsigned int, l);
    }
}

// ---------------------------------------------------------------- VALU kernel
// Weights stream: [group][n_stream][2*G] floats (Re, Im per rhs); positions
// in the per-step layout.  EPI_COMBINE sums conj(S_c) * acc_c over the rhs.
enum { EPI_COMBINE = 0, EPI_PERCOIL = 1 };

template <int G, bool D3, int EPI>
__global__ __launch_bounds__(NU_THREADS) void k_nudft_valu(
    const float4 *__restrict__ lanes, int n_lane, const double *__restrict__ spos,
    const float *__restrict__ wstream, int n_stream, int chunk_len, float sign, int nrhs,
    const float *__restrict__ s_r, const float *__restrict__ s_i,
    float2 *__restrict__ partial) {
    const int a = blockIdx.x * NU_THREADS + threadIdx.x;
    const int chunk = blockIdx.y;
    const int group = blockIdx.z;
    const int nchunks = gridDim.y;
    const int b0 = chunk * chunk_len;
    const int b1 = min(n_stream, b0 + chunk_len);
    const float *__restrict__ gw = wstream + (size_t)group * n_stream * (2 * G);

    const float4 Pf = lanes[min(a, n_lane - 1)];
    const double px = Pf.x, py = Pf.y, pz = Pf.z, pw = Pf.w;

    float hr[G], hi[G], lr[G], li[G];
#pragma unroll
    for (int g = 0; g < G; ++g) { hr[g] = hi[g] = lr[g] = li[g] = 0.f; }

    for (int b = b0; b < b1; b += NU_BLK) {
        const int be = min(b + NU_BLK, b1);
        float ar[G], ai[G];
#pragma unroll
        for (int g = 0; g < G; ++g) { ar[g] = ai[g] = 0.f; }
        for (int j = b; j < be; ++j) {
            const double *B = spos + (size_t)(j / MF_K) * pos_stride(D3) + (j % MF_K);
            double u = B[2 * MF_K] * pw;
            if constexpr (D3) u = __builtin_fma(B[3 * MF_K], pz, u);
            u = __builtin_fma(B[MF_K], py, u);
            u = __builtin_fma(B[0], px, u);
            const float r = rev_fract(u);
            const float cs = __builtin_amdgcn_cosf(r);
            const float sn = sign * __builtin_amdgcn_sinf(r);
            const float *__restrict__ e = gw + (size_t)j * (2 * G);
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const float wr = e[2 * g], wi = e[2 * g + 1];
                ar[g] = __builtin_fmaf(wr, cs, __builtin_fmaf(-wi, sn, ar[g]));
                ai[g] = __builtin_fmaf(wr, sn, __builtin_fmaf(wi, cs, ai[g]));
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
        float orr = 0.f, oi = 0.f;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int c = group * G + g;
            if (c < nrhs) {
                const float vr = hr[g] + lr[g], vi = hi[g] + li[g];
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
            if (c < nrhs)
                partial[((size_t)chunk * nrhs + c) * n_lane + a] =
                    make_float2(hr[g] + lr[g], hi[g] + li[g]);
        }
    }
}

// ---------------------------------------------------------------- MFMA kernel
// One wave owns 64 points (M tiles A = points 0-31, B = points 32-63) and
// TILES 32-column N tiles (16 rhs x {Re, Im} each).  Per step of 8 samples
// lane l evaluates the phases of point l and splits cos and sin into fp16
// hi/lo vectors; v_permlane32_swap(cos, sin) then yields
// [cos of A | sin of A] and [cos of B | sin of B], i.e. the A operands with
// K rows (Re|Im, sample) that k_pack_mfma's B operands expect.  B operands
// are fetched two steps ahead into two alternating register sets (bpack
// holds MF_BPAD zero steps after the last).  The single-job launch
// software-pipelines the steps: the MFMAs of step s are interleaved with the
// phase, sin/cos and split work of step s + 1 (one MFMA, then MF_PIPE_VALU
// VALU instructions), so each wave keeps the matrix core and the VALU
// issuing together.  tot stays in registers and the two-sum
// compensation in LDS.  Output planes: partial[chunk][TILES*32][n_lane]
// (plane = tile*32 + part*16 + rhs%16).  `one` is 1.0f, passed at run time
// so the lo split compiles to v_fma_mix.
template <int TILES>
struct MfmaLds {
    float cmp[NU_THREADS / 64][2 * TILES * 16][64];
};

// Workgroup (bx, chunk) of one NUDFT job; shared by the single-job and the
// batched launch.
template <int TILES, bool D3, bool PAIRS>
__device__ __forceinline__ void mfma_body(const float4 *__restrict__ lanes, int n_lane,
                                          const double *__restrict__ spos,
                                          const BPair *__restrict__ bpack, int n_steps,
                                          int steps_per_chunk, float *__restrict__ partial,
                                          float one, int bx, int chunk, MfmaLds<TILES> &lds) {
    constexpr int NA = 2 * TILES;  // accumulators [m tile][n tile]
    constexpr int WAVES = NU_THREADS / 64;
    auto &cmp_lds = lds.cmp;
    const int l = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wt = bx * WAVES + wave;
    const int h = l >> 5;
    const float4 Pf = lanes[min(wt * MF_WAVE_PTS + l, n_lane - 1)];
    const double px = Pf.x, py = Pf.y, pz = Pf.z, pw = Pf.w;
    const int s0 = chunk * steps_per_chunk;
    const int s1 = min(n_steps, s0 + steps_per_chunk);

    floatx16 acc[NA];
    float tot[NA][16];
#pragma unroll
    for (int q = 0; q < NA; ++q) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            acc[q][i] = 0.f;
            tot[q][i] = 0.f;
            cmp_lds[wave][q * 16 + i][l] = 0.f;
        }
    }
    // B operands in two register sets that alternate by step: step st uses
    // b0 (even offset in a pair) or b1, and the set is refilled with step
    // st + 2 as soon as st's MFMAs have read it, so no register holding a
    // load in flight is ever copied (bpack holds MF_BPAD zero steps after the
    // last, so st + 2 and st + 3 stay inside it)
    BPair b0[TILES], b1[TILES];
    const int sb = min(s0, n_steps);
#pragma unroll
    for (int t = 0; t < TILES; ++t) {
        b0[t] = bpack[((size_t)sb * TILES + t) * 64 + l];
        b1[t] = bpack[((size_t)(sb + 1) * TILES + t) * 64 + l];
    }
    // positions of the current step (wave-uniform, scalar loads); the next
    // step's are loaded as soon as these are consumed, so their latency
    // overlaps the MFMAs instead of stalling the next step's phases
    constexpr int NPOS = pos_stride(D3);
    double pos[NPOS];
    {
        const cdouble_ptr B = (cdouble_ptr)spos + (size_t)min(s0, max(n_steps - 1, 0)) * NPOS;
#pragma unroll
        for (int j = 0; j < NPOS; ++j) pos[j] = B[j];
    }

    struct AOp {
        half8 hA, hB, lA, lB;
    };
    // A operands of the step whose positions are in pos; pos is then
    // reloaded with step `next` (clamped inside the chunk)
    auto operands = [&](int next, AOp &o) {
        // written stage by stage over the 8 samples, so that consecutive
        // VALU instructions are independent (1.5% faster than sample by
        // sample, where each phase chain issued back to back)
        float cv[MF_K], sv[MF_K];
        double u[MF_K];
        float r[MF_K];
#pragma unroll
        for (int j = 0; j < MF_K; ++j) u[j] = pos[2 * MF_K + j] * pw;
        if constexpr (D3) {
#pragma unroll
            for (int j = 0; j < MF_K; ++j) u[j] = __builtin_fma(pos[3 * MF_K + j], pz, u[j]);
        }
#pragma unroll
        for (int j = 0; j < MF_K; ++j) u[j] = __builtin_fma(pos[MF_K + j], py, u[j]);
#pragma unroll
        for (int j = 0; j < MF_K; ++j) u[j] = __builtin_fma(pos[j], px, u[j]);
#pragma unroll
        for (int j = 0; j < MF_K; ++j) r[j] = rev_fract(u[j]);
#pragma unroll
        for (int j = 0; j < MF_K; ++j) cv[j] = __builtin_amdgcn_cosf(r[j]);
#pragma unroll
        for (int j = 0; j < MF_K; ++j) sv[j] = __builtin_amdgcn_sinf(r[j]);
        {
            const cdouble_ptr B = (cdouble_ptr)spos + (size_t)min(next, s1 - 1) * NPOS;
#pragma unroll
            for (int j = 0; j < NPOS; ++j) pos[j] = B[j];
        }
        u32x4 chu, clu, shu, slu;
        split8(cv, one, chu, clu);
        split8(sv, one, shu, slu);
        u32x4 hA, hB, lA, lB;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const auto rh = __builtin_amdgcn_permlane32_swap(chu[i], shu[i], false, false);
            hA[i] = rh[0];
            hB[i] = rh[1];
            const auto rl = __builtin_amdgcn_permlane32_swap(clu[i], slu[i], false, false);
            lA[i] = rl[0];
            lB[i] = rl[1];
        }
        o.hA = __builtin_bit_cast(half8, hA);
        o.hB = __builtin_bit_cast(half8, hB);
        o.lA = __builtin_bit_cast(half8, lA);
        o.lB = __builtin_bit_cast(half8, lB);
    };
    auto mfmas = [&](const AOp &o, const BPair(&b)[TILES]) {
#pragma unroll
        for (int t = 0; t < TILES; ++t) {
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(o.hA, b[t].hi, acc[t], 0, 0, 0);
            acc[TILES + t] =
                __builtin_amdgcn_mfma_f32_32x32x16_f16(o.hB, b[t].hi, acc[TILES + t], 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < TILES; ++t) {
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(o.hA, b[t].lo, acc[t], 0, 0, 0);
            acc[TILES + t] =
                __builtin_amdgcn_mfma_f32_32x32x16_f16(o.hB, b[t].lo, acc[TILES + t], 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < TILES; ++t) {
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(o.lA, b[t].hi, acc[t], 0, 0, 0);
            acc[TILES + t] =
                __builtin_amdgcn_mfma_f32_32x32x16_f16(o.lB, b[t].hi, acc[TILES + t], 0, 0, 0);
        }
    };
    // step st: its MFMAs (operands cur, B operands b) interleaved with the
    // VALU work of step st + 1 (into nxt)
    auto pstep = [&](int st, const AOp &cur, AOp &nxt, const BPair(&b)[TILES]) {
        operands(st + 2, nxt);
        mfmas(cur, b);
#pragma unroll
        for (int k = 0; k < 6 * TILES; ++k) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, MF_PIPE_VALU, 0);
        }
    };
    auto step = [&](int st, const BPair(&b)[TILES]) {
        AOp o;
        operands(st + 1, o);
        mfmas(o, b);
    };
    auto refill = [&](BPair(&b)[TILES], int st) {
#pragma unroll
        for (int t = 0; t < TILES; ++t) b[t] = bpack[((size_t)st * TILES + t) * 64 + l];
    };

    AOp op0, op1;
    if constexpr (PAIRS) {
        // prologue: operands of s0; pos then holds step s0 + 1
        if (s0 < s1) operands(s0 + 1, op0);
    }
    for (int f0 = s0; f0 < s1; f0 += MF_FOLD) {
        const int f1 = min(s1, f0 + MF_FOLD);
        if constexpr (PAIRS) {
            int st = f0;
            // invariant at the top: op0 and b0 hold step st, b1 step st + 1,
            // pos step st + 1
            for (; st + 1 < f1; st += 2) {
                pstep(st, op0, op1, b0);
                refill(b0, st + 2);
                pstep(st + 1, op1, op0, b1);
                refill(b1, st + 3);
            }
            if (st < f1) {   // odd tail of a chunk: one step, then shift the sets
                pstep(st, op0, op1, b0);
                op0 = op1;
#pragma unroll
                for (int t = 0; t < TILES; ++t) b0[t] = b1[t];
                refill(b1, st + 2);
            }
        } else {
            // one set ahead, copied per step: fewer registers (the batched
            // launch, whose job lookup leaves less room)
            for (int st = f0; st < f1; ++st) {
                BPair cur[TILES];
#pragma unroll
                for (int t = 0; t < TILES; ++t) {
                    cur[t] = b0[t];
                    b0[t] = bpack[((size_t)(st + 1) * TILES + t) * 64 + l];
                }
                step(st, cur);
            }
        }
#pragma unroll
        for (int q = 0; q < NA; ++q) {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                float sum, err;
                two_sum(tot[q][i], acc[q][i], sum, err);
                tot[q][i] = sum;
                cmp_lds[wave][q * 16 + i][l] += err;
                acc[q][i] = 0.f;
            }
        }
    }
    // D layout of 32x32 MFMA: lane l, register i -> column l%32,
    // row 8*(i/4) + 4*(l/32) + i%4
    const int n = l & 31;
    constexpr int nplanes = TILES * 32;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
#pragma unroll
        for (int t = 0; t < TILES; ++t) {
            const int q = m * TILES + t;
            float *__restrict__ plane = partial + ((size_t)chunk * nplanes + t * 32 + n) * n_lane;
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int ap = wt * MF_WAVE_PTS + m * 32 + 8 * (i >> 2) + 4 * h + (i & 3);
                if (ap < n_lane) plane[ap] = tot[q][i] + cmp_lds[wave][q * 16 + i][l];
            }
        }
    }
}

template <int TILES, bool D3>
__global__ __launch_bounds__(NU_THREADS) __attribute__((amdgpu_waves_per_eu(TILES == 1 ? 3 : 2)))
void k_nudft_mfma(const float4 *__restrict__ lanes, int n_lane, const double *__restrict__ spos,
                  const BPair *__restrict__ bpack, int n_steps, int steps_per_chunk,
                  float *__restrict__ partial, float one) {
    __shared__ MfmaLds<TILES> lds;
    mfma_body<TILES, D3, true>(lanes, n_lane, spos, bpack, n_steps, steps_per_chunk, partial,
                               one, blockIdx.x, blockIdx.y, lds);
}

// A wave-uniform pointer moved into scalar registers.
__device__ __forceinline__ const void *uniform_ptr(const void *p) {
    const unsigned long long v = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return (const void *)(((unsigned long long)hi << 32) | lo);
}

// One job of a batched launch (images of different sizes in one grid).
struct MfmaJob {
    const float4 *lanes;
    const double *spos;
    const BPair *bpack;
    float *partial;
    int n_lane, n_steps, steps_per_chunk;
    int wg_begin;   // first flattened workgroup of this job
    int wgs_x;      // workgroups along the points; chunks = count / wgs_x
    int pad[3];
};

// Persistent-grid batch: flattened workgroup -> (job, bx, chunk); the job
// table is scanned with wave-uniform (scalar) loads.
template <int TILES, bool D3>
__global__ __launch_bounds__(NU_THREADS) __attribute__((amdgpu_waves_per_eu(TILES == 1 ? 3 : 2)))
void k_nudft_mfma_batch(const MfmaJob *__restrict__ jobs, int njobs, float one) {
    __shared__ MfmaLds<TILES> lds;
    const int g = blockIdx.x;
    int j = 0;
    while (j + 1 < njobs && jobs[j + 1].wg_begin <= g) ++j;
    // the job's fields are the same for every lane: held in scalar
    // registers (loaded per lane, they would occupy vector registers and the
    // step loop would spill)
    const MfmaJob &Jm = jobs[j];
    const float4 *lanes = (const float4 *)uniform_ptr(Jm.lanes);
    const double *spos = (const double *)uniform_ptr(Jm.spos);
    const BPair *bpack = (const BPair *)uniform_ptr(Jm.bpack);
    float *partial = (float *)uniform_ptr(Jm.partial);
    const int n_lane = __builtin_amdgcn_readfirstlane(Jm.n_lane);
    const int n_steps = __builtin_amdgcn_readfirstlane(Jm.n_steps);
    const int spc = __builtin_amdgcn_readfirstlane(Jm.steps_per_chunk);
    const int wgs_x = __builtin_amdgcn_readfirstlane(Jm.wgs_x);
    const int rem = g - __builtin_amdgcn_readfirstlane(Jm.wg_begin);
    mfma_body<TILES, D3, true>(lanes, n_lane, spos, bpack, n_steps, spc, partial, one,
                               rem % wgs_x, rem / wgs_x, lds);
}

// ---- packing ---------------------------------------------------------------

// stream positions as FP64 in the per-step layout [x8][y8][w8][z8 if 3D],
// zero-padded to a multiple of MF_K samples
__global__ void k_spos(int n, int n_pad, const float *x, const float *y, const float *z,
                       const float *w, double wscale, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pad) return;
    const bool d3 = z != nullptr;
    double *o = out + (size_t)(i / MF_K) * pos_stride(d3) + (i % MF_K);
    const bool in = i < n;
    o[0] = in ? (double)x[i] : 0.0;
    o[MF_K] = in ? (double)y[i] : 0.0;
    o[2 * MF_K] = in && w ? (double)w[i] * wscale : 0.0;
    if (d3) o[3 * MF_K] = in ? (double)z[i] : 0.0;
}

__global__ void k_lanes_from_arrays(int n, const float *x, const float *y, const float *z,
                                    const float *w, double wscale, float4 *lanes) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    lanes[i] = make_float4(x[i], y[i], z ? z[i] : 0.f,
                           w ? (float)((double)w[i] * wscale) : 0.f);
}

// Half of the standard Q grid.  Q(-r) = conj(Q(r)) because the Q weights are
// real, so only rows y <= 0 (grid indices [0, n1)) and the points of rows
// y > 0 whose mirror falls off the grid (x = -Nx, or z = -Nz in 3D; ne_row
// per row) are evaluated; k_q_fill mirrors the rest.
__device__ __forceinline__ void qhalf_edge(int e, int Ny, int M2, int M3, int ne_row, int &iy,
                                           int &ix, int &iz) {
    iy = Ny + 1 + e / ne_row;
    const int r = e % ne_row;
    if (M3 > 1) {
        ix = r < M2 ? r : 0;
        iz = r < M2 ? 0 : r - M2 + 1;
    } else {
        ix = 0;
        iz = 0;
    }
}

__global__ void k_lanes_qhalf(int Nx, int Ny, int Nz, int n1, int ne_row, int n_comp,
                              float4 *lanes) {
    const int M2 = 2 * Nx, M3 = Nz > 1 ? 2 * Nz : 1;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_comp) return;
    int iy, ix, iz;
    if (j < n1) {
        iz = j % M3;
        ix = (j / M3) % M2;
        iy = j / (M3 * M2);
    } else {
        qhalf_edge(j - n1, Ny, M2, M3, ne_row, iy, ix, iz);
    }
    lanes[j] = make_float4((float)(ix - Nx) / (float)Nx, (float)(iy - Ny) / (float)Ny,
                           Nz > 1 ? (float)(iz - Nz) / (float)Nz : 0.f, 0.f);
}

// half-grid values [l][n_comp] -> full grid [l][nq]
__global__ void k_q_fill(int Nx, int Ny, int Nz, int n1, int ne_row, int n_comp, int nseg,
                         const float *hr, const float *hi, float *q_r, float *q_i) {
    const int M2 = 2 * Nx, M3 = Nz > 1 ? 2 * Nz : 1;
    const long nq = 2L * Ny * M2 * M3;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nq) return;
    const int iz = (int)(g % M3);
    const int ix = (int)((g / M3) % M2);
    const int iy = (int)(g / ((long)M3 * M2));
    long src;
    float sgn = 1.f;
    if (iy <= Ny) {
        src = g;
    } else if (ix == 0 || (M3 > 1 && iz == 0)) {
        const int r = M3 > 1 ? (iz == 0 ? ix : M2 + iz - 1) : 0;
        src = n1 + (long)(iy - Ny - 1) * ne_row + r;
    } else {
        const int my = 2 * Ny - iy, mx = 2 * Nx - ix, mz = M3 > 1 ? 2 * Nz - iz : 0;
        src = ((long)my * M2 + mx) * M3 + mz;
        sgn = -1.f;
    }
    for (int l = 0; l < nseg; ++l) {
        q_r[(size_t)l * nq + g] = hr[(size_t)l * n_comp + src];
        q_i[(size_t)l * nq + g] = sgn * hi[(size_t)l * n_comp + src];
    }
}

// Weights W[c][b] (complex, rhs-major) for the three operators.
// adjoint: conj(phi) d (ComputeRhoPhiGPU, computeFH.cmem.cu:112-127)
__global__ void k_w_adjoint(int num_k, int ncoils, const float *phi_r, const float *phi_i,
                            const float *d_r, const float *d_i, float2 *W) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)num_k * ncoils) return;
    const float dr = d_r[i], di = d_i[i];
    if (phi_r) {
        const float pr = phi_r[i], pi = phi_i[i];
        W[i] = make_float2(pr * dr + pi * di, pr * di - pi * dr);
    } else {
        W[i] = make_float2(dr, di);
    }
}

// forward: S_c rho
__global__ void k_w_forward(int num_x, int ncoils, const float *rho_r, const float *rho_i,
                            const float *s_r, const float *s_i, float2 *W) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)num_x * ncoils) return;
    const int v = (int)(i % num_x);
    const float rr = rho_r[v], ri = rho_i[v];
    float sr = 1.f, si = 0.f;
    if (s_r) {
        sr = s_r[i];
        si = s_i[i];
    }
    W[i] = make_float2(sr * rr - si * ri, sr * ri + si * rr);
}

// forward of a stack: W[(c P + p)][r] = S_c(r) rho(r) e^{-i 2pi kz_p z_r}
// (the kz factor of partition p, FP64 phase)
__global__ void k_w_forward_stack(int num_x, int ncoils, int P, int K, const float *kz,
                                  const float *z, const float *rho_r, const float *rho_i,
                                  const float *s_r, const float *s_i, float2 *W) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)num_x * ncoils * P) return;
    const int v = (int)(i % num_x);
    const int q = (int)(i / num_x);
    const int c = q / P, p = q % P;
    const double rr = rho_r[v], ri = rho_i[v];
    double sr = 1.0, si = 0.0;
    if (s_r) {
        sr = s_r[(size_t)c * num_x + v];
        si = s_i[(size_t)c * num_x + v];
    }
    const double ar = sr * rr - si * ri, ai = sr * ri + si * rr;
    double sn, cs;
    sincospi(2.0 * ((double)kz[(size_t)p * K] * (double)z[v]), &sn, &cs);
    W[i] = make_float2((float)(ar * cs + ai * sn), (float)(ai * cs - ar * sn));
}

// Q: a_l(t_k) |phi_k|^2 with the Hanning interpolator of hanning_d
// (computeQ.cmem.cu:125-138), evaluated in FP64.
__global__ void k_w_q(int num_k, int nseg, const float *t, const float *phi_r,
                      const float *phi_i, float tau, float t0, float2 *W) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)num_k * nseg) return;
    const int l = (int)(i / num_k), k = (int)(i % num_k);
    double mag = 1.0;
    if (phi_r) {
        const double pr = phi_r[k], pi = phi_i[k];
        mag = pr * pr + pi * pi;
    }
    const double arg = (double)t[k] - (double)tau * l - (double)t0;
    double w = 0.0;
    if (fabs(arg) < (double)tau) w = 0.5 + 0.5 * cos(M_PI * arg / (double)tau);
    W[i] = make_float2((float)(w * mag), 0.f);
}

// VALU stream: [group][n][2G] from W[c][n]
__global__ void k_pack_valu(int n, int nrhs, int G, const float2 *W, float *out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const int group = blockIdx.y;
    if (b >= n) return;
    float *e = out + ((size_t)group * n + b) * (2 * G);
    for (int g = 0; g < G; ++g) {
        const int c = group * G + g;
        const float2 w = c < nrhs ? W[(size_t)c * n + b] : make_float2(0.f, 0.f);
        e[2 * g] = w.x;
        e[2 * g + 1] = w.y;
    }
}

// max |Re|, |Im| over W (non-negative floats compare as unsigned ints)
__global__ void k_absmax(const float2 *W, size_t n, unsigned int *out) {
    float m = 0.f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const float2 w = W[i];
        m = fmaxf(m, fmaxf(fabsf(w.x), fabsf(w.y)));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

__device__ __forceinline__ int w_shift(unsigned int maxbits) {
    // exponent s with max|W| * 2^s < 2^W_EXP
    const float m = __uint_as_float(maxbits);
    if (!(m > 0.f)) return 0;
    int e;
    frexpf(m, &e);  // m = f * 2^e, f in [0.5, 1)
    return W_EXP - e;
}

// B operands for k_nudft_mfma: step st (8 samples), tile t, lane l:
// K row (h, j) = (Re|Im of E, sample 8*st + j); column n = l%32 -> part n/16,
// rhs t*16 + n%16.  Coefficients: out_r = Er Wr - sgn Ei Wi,
// out_i = Er Wi + sgn Ei Wr.
__global__ void k_pack_mfma(int n, int nrhs, int tiles, int n_steps, float sign,
                            const float2 *W, const unsigned int *maxbits, BPair *out) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)n_steps * tiles * 64) return;
    const int l = (int)(idx & 63);
    const int t = (int)((idx >> 6) % tiles);
    const int st = (int)((idx >> 6) / tiles);
    const int h = l >> 5, nn = l & 31;
    const int part = nn >> 4;
    const int c = t * MF_COLS + (nn & 15);
    const float scale = ldexpf(1.f, w_shift(*maxbits));
    BPair bp;
#pragma unroll
    for (int j = 0; j < MF_K; ++j) {
        const int b = st * MF_K + j;
        float v = 0.f;
        if (b < n && c < nrhs) {
            const float2 w = W[(size_t)c * n + b];
            if (h == 0) v = part == 0 ? w.x : w.y;
            else v = part == 0 ? -sign * w.y : sign * w.x;
        }
        v *= scale;
        const _Float16 vh = (_Float16)v;
        bp.hi[j] = vh;
        bp.lo[j] = (_Float16)(v - (float)vh);
    }
    out[idx] = bp;
}

// ---- reductions over chunk partials (FP64) ---------------------------------

__global__ void k_reduce_combine(const float2 *partial, int nparts, int n, float *out_r,
                                 float *out_i) {
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

// VALU per-rhs partials [chunk][c][n] float2 -> out[c][n]; cumulative: prefix over c
__global__ void k_reduce_percoil(const float2 *partial, int nchunks, int nrhs, int n,
                                 int cumulative, float *out_r, float *out_i) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    double cr = 0.0, ci = 0.0;
    for (int c = 0; c < nrhs; ++c) {
        double sr = 0.0, si = 0.0;
        for (int p = 0; p < nchunks; ++p) {
            const float2 v = partial[((size_t)p * nrhs + c) * n + a];
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

// MFMA planes [chunk][tiles*32][n] -> combine with conj(S_c) (mode 0),
// per-rhs output (mode 1), per-rhs prefix sums (mode 2); undoes the W scaling.
__global__ void k_reduce_mfma(const float *partial, int nchunks, int tiles, int nrhs, int n,
                              int mode, const unsigned int *maxbits, const float *s_r,
                              const float *s_i, float *out_r, float *out_i) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const double unscale = ldexp(1.0, -w_shift(*maxbits));
    const int nplanes = tiles * 32;
    double orr = 0.0, oi = 0.0, cr = 0.0, ci = 0.0;
    for (int c = 0; c < nrhs; ++c) {
        const int t = c / MF_COLS, cc = c % MF_COLS;
        double vr = 0.0, vi = 0.0;
        for (int p = 0; p < nchunks; ++p) {
            const size_t base = (size_t)p * nplanes + t * 32;
            vr += partial[(base + cc) * n + a];
            vi += partial[(base + MF_COLS + cc) * n + a];
        }
        vr *= unscale;
        vi *= unscale;
        if (mode == 0) {
            double sr = 1.0, si = 0.0;
            if (s_r) {
                sr = s_r[(size_t)c * n + a];
                si = s_i[(size_t)c * n + a];
            }
            orr += sr * vr + si * vi;
            oi += sr * vi - si * vr;
        } else {
            if (mode == 2) {
                cr += vr;
                ci += vi;
                vr = cr;
                vi = ci;
            }
            out_r[(size_t)c * n + a] = (float)vr;
            out_i[(size_t)c * n + a] = (float)vi;
        }
    }
    if (mode == 0) {
        out_r[a] = (float)orr;
        out_i[a] = (float)oi;
    }
}

// Per-rhs output (mode 1) when there are few points and many chunks: one
// workgroup per (64 points, rhs), its four waves split the chunks and meet
// in LDS (FP64).
__global__ __launch_bounds__(256) void k_reduce_mfma_rhs(const float *partial, int nchunks,
                                                         int tiles, int n,
                                                         const unsigned int *maxbits,
                                                         float *out_r, float *out_i) {
    __shared__ double red[2][4][64];
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int a = blockIdx.x * 64 + l;
    const int c = blockIdx.y;
    const int t = c / MF_COLS, cc = c % MF_COLS;
    const int nplanes = tiles * 32;
    double vr = 0.0, vi = 0.0;
    if (a < n) {
        for (int p = w; p < nchunks; p += 4) {
            const size_t base = (size_t)p * nplanes + t * 32;
            vr += partial[(base + cc) * n + a];
            vi += partial[(base + MF_COLS + cc) * n + a];
        }
    }
    red[0][w][l] = vr;
    red[1][w][l] = vi;
    __syncthreads();
    if (w == 0 && a < n) {
        const double unscale = ldexp(1.0, -w_shift(*maxbits));
        const double sr = red[0][0][l] + red[0][1][l] + red[0][2][l] + red[0][3][l];
        const double si = red[1][0][l] + red[1][1][l] + red[1][2][l] + red[1][3][l];
        out_r[(size_t)c * n + a] = (float)(sr * unscale);
        out_i[(size_t)c * n + a] = (float)(si * unscale);
    }
}

// ---- dispatch ----------------------------------------------------------------

using VFn = void (*)(const float4 *, int, const double *, const float *, int, int, float, int,
                     const float *, const float *, float2 *);

template <int G, int EPI>
VFn pick_valu_d3(bool d3) {
    return d3 ? (VFn)k_nudft_valu<G, true, EPI> : (VFn)k_nudft_valu<G, false, EPI>;
}

template <int EPI>
VFn pick_valu(int G, bool d3) {
    switch (G) {
        case 1: return pick_valu_d3<1, EPI>(d3);
        case 2: return pick_valu_d3<2, EPI>(d3);
        default: return pick_valu_d3<4, EPI>(d3);
    }
}

using MFn = void (*)(const float4 *, int, const double *, const BPair *, int, int, float *, float);

MFn pick_mfma(int tiles, bool d3) {
    if (tiles == 1) return d3 ? (MFn)k_nudft_mfma<1, true> : (MFn)k_nudft_mfma<1, false>;
    return d3 ? (MFn)k_nudft_mfma<2, true> : (MFn)k_nudft_mfma<2, false>;
}

// right-hand sides at or above which the MFMA kernel is used
constexpr int MFMA_MIN_RHS = 3;

// Chunks so the launch holds enough workgroups to fill the GPU.
int pick_chunks(imp_ctx *ctx, long wgs, int n_units, int min_units_per_chunk) {
    const long target = 8L * ctx->num_cus;
    int chunks = 1;
    while (wgs * chunks < target && n_units / (chunks * 2) >= min_units_per_chunk) chunks *= 2;
    return chunks;
}

enum class Out { Combine, PerRhs, PerRhsCumulative };

// Shared driver: lanes + stream positions + W are built by the caller.
struct Job {
    const float4 *lanes;
    int n_lane;
    const double *spos;      // per-step layout, padded to a multiple of MF_K
    int n_stream;
    const float2 *W;         // [nrhs][n_stream]
    int nrhs;
    bool d3;
    float sign;
    Out out;
    const float *s_r, *s_i;  // Combine only
    float *out_r, *out_i;
    const char *prof = "k_nudft_mfma";   // kernel-timer name of the MFMA launch
};

// Where one call's device scratch comes from: the context's growable
// workspaces (base == nullptr), or a fixed region carved front to back in
// 256-byte units (the allocation-free entry point).
struct Scratch {
    imp_ctx *ctx;
    char *base = nullptr;
    size_t cap = 0, used = 0;
    static size_t unit(size_t bytes) { return (bytes + 255) & ~(size_t)255; }
    int get(int slot, size_t bytes, void **out) {
        if (!base) return workspace(ctx, slot, bytes, out);
        if (unit(bytes) > cap - used)
            return fail(IMP_EUNSUPPORTED, "NUDFT: the fixed scratch region (%zu bytes) is %zu "
                                          "bytes short", cap, unit(bytes) - (cap - used));
        *out = base + used;
        used += unit(bytes);
        return IMP_OK;
    }
    // chunks whose partials still fit (fixed region), at most `want`
    int fit_chunks(int want, size_t per_chunk) const {
        if (!base) return want;
        const size_t n = (cap - used) / unit(per_chunk);
        return (int)std::min<size_t>((size_t)want, n);
    }
    // the kernel timer records HIP events; the fixed-region path records none
    int prof(const char *name) const { return base ? -1 : prof_begin(ctx, name); }
};

__global__ void k_zero_u32(unsigned int *p) { *p = 0u; }

int run_job(imp_ctx *ctx, const Job &J, Scratch &S) {
    const hipStream_t st = ctx->stream;
    const int n = J.n_lane;
    if (J.nrhs >= MFMA_MIN_RHS) {
        int total_tiles = ceil_div(J.nrhs, MF_COLS);
        if (total_tiles > 2)
            return fail(IMP_EUNSUPPORTED, "NUDFT: %d right-hand sides (at most 32)", J.nrhs);
        const int tiles = total_tiles;
        const int n_steps = ceil_div(J.n_stream, MF_K);
        // max |W| for the power-of-two scale (cleared by a kernel, not a
        // memset node, so the sequence can be captured in a graph)
        unsigned int *maxbits = nullptr;
        void *mb = nullptr, *bp = nullptr, *part = nullptr;
        IMP_TRY(S.get(WS_MAXBITS, 256, &mb));
        maxbits = (unsigned int *)mb;
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, st, maxbits);
        hipLaunchKernelGGL(k_absmax, dim3(512), dim3(256), 0, st, J.W,
                           (size_t)J.nrhs * J.n_stream, maxbits);
        IMP_CHECK_LAUNCH();
        // MF_BPAD zero steps after the last: the kernel fetches up to step st + 2
        const size_t nb = (size_t)(n_steps + MF_BPAD) * tiles * 64;
        IMP_TRY(S.get(WS_BPACK, nb * sizeof(BPair), &bp));
        hipLaunchKernelGGL(k_pack_mfma, dim3(ceil_div((long)nb, 256)), dim3(256), 0, st,
                           J.n_stream, J.nrhs, tiles, n_steps + MF_BPAD, J.sign, J.W, maxbits,
                           (BPair *)bp);
        IMP_CHECK_LAUNCH();
        const long wgs = ceil_div(n, MF_WAVE_PTS * (NU_THREADS / 64));
        const size_t per_chunk = (size_t)tiles * 32 * n * sizeof(float);
        const int chunks = S.fit_chunks(pick_chunks(ctx, wgs, n_steps, 4 * MF_FOLD), per_chunk);
        if (chunks < 1) return fail(IMP_EUNSUPPORTED, "NUDFT: no room for one chunk's partials");
        const int spc = ceil_div(n_steps, chunks);
        IMP_TRY(S.get(WS_PARTIAL, (size_t)chunks * per_chunk, &part));
        const int rec = S.prof(J.prof);
        hipLaunchKernelGGL(pick_mfma(tiles, J.d3), dim3((unsigned)wgs, chunks), dim3(NU_THREADS),
                           0, st, J.lanes, n, J.spos, (const BPair *)bp, n_steps, spc,
                           (float *)part, 1.0f);
        prof_end(ctx, rec);
        IMP_CHECK_LAUNCH();
        const int mode = J.out == Out::Combine ? 0 : (J.out == Out::PerRhs ? 1 : 2);
        if (mode == 1 && (long)n * J.nrhs < 64L * ctx->num_cus * 16)
            hipLaunchKernelGGL(k_reduce_mfma_rhs, dim3(ceil_div(n, 64), J.nrhs), dim3(256), 0, st,
                               (const float *)part, chunks, tiles, n, maxbits, J.out_r, J.out_i);
        else
            hipLaunchKernelGGL(k_reduce_mfma, dim3(ceil_div(n, 256)), dim3(256), 0, st,
                               (const float *)part, chunks, tiles, J.nrhs, n, mode, maxbits,
                               J.s_r, J.s_i, J.out_r, J.out_i);
        IMP_CHECK_LAUNCH();
        return IMP_OK;
    }
    // VALU path
    const int G = J.nrhs <= 1 ? 1 : (J.nrhs <= 2 ? 2 : 4);
    const int groups = ceil_div(J.nrhs, G);
    void *ws = nullptr, *part = nullptr;
    IMP_TRY(S.get(WS_BPACK, (size_t)groups * J.n_stream * 2 * G * sizeof(float), &ws));
    hipLaunchKernelGGL(k_pack_valu, dim3(ceil_div(J.n_stream, 256), groups), dim3(256), 0, st,
                       J.n_stream, J.nrhs, G, J.W, (float *)ws);
    IMP_CHECK_LAUNCH();
    const int lane_blocks = ceil_div(n, NU_THREADS);
    const bool combine = J.out == Out::Combine;
    const size_t per_chunk = (combine ? (size_t)groups : (size_t)J.nrhs) * n * sizeof(float2);
    const int chunks = S.fit_chunks(
        pick_chunks(ctx, (long)lane_blocks * groups, J.n_stream, 4 * NU_BLK), per_chunk);
    if (chunks < 1) return fail(IMP_EUNSUPPORTED, "NUDFT: no room for one chunk's partials");
    int chunk_len = ceil_div(ceil_div(J.n_stream, chunks), NU_BLK) * NU_BLK;
    IMP_TRY(S.get(WS_PARTIAL, (size_t)chunks * per_chunk, &part));
    VFn fn = combine ? pick_valu<EPI_COMBINE>(G, J.d3) : pick_valu<EPI_PERCOIL>(G, J.d3);
    const int rec = S.prof("k_nudft_valu");
    hipLaunchKernelGGL(fn, dim3(lane_blocks, chunks, groups), dim3(NU_THREADS), 0, st, J.lanes,
                       n, J.spos, (const float *)ws, J.n_stream, chunk_len, J.sign, J.nrhs,
                       J.s_r, J.s_i, (float2 *)part);
    prof_end(ctx, rec);
    IMP_CHECK_LAUNCH();
    if (combine) {
        hipLaunchKernelGGL(k_reduce_combine, dim3(ceil_div(n, 256)), dim3(256), 0, st,
                           (const float2 *)part, groups * chunks, n, J.out_r, J.out_i);
    } else {
        hipLaunchKernelGGL(k_reduce_percoil, dim3(ceil_div(n, 256)), dim3(256), 0, st,
                           (const float2 *)part, chunks, J.nrhs, n,
                           J.out == Out::PerRhsCumulative ? 1 : 0, J.out_r, J.out_i);
    }
    IMP_CHECK_LAUNCH();
    return IMP_OK;
}

int run_job(imp_ctx *ctx, const Job &J) {
    Scratch S{ctx};
    return run_job(ctx, J, S);
}

size_t spos_bytes(int n, bool d3) {
    return (size_t)(ceil_div(n, MF_K)) * pos_stride(d3) * sizeof(double);
}

int build_spos(imp_ctx *ctx, Scratch &S, int n, const float *x, const float *y, const float *z,
               const float *w, double wscale, const double **out) {
    void *p = nullptr;
    const int n_pad = ceil_div(n, MF_K) * MF_K;
    IMP_TRY(S.get(WS_SPOS, spos_bytes(n, z != nullptr), &p));
    hipLaunchKernelGGL(k_spos, dim3(ceil_div(n_pad, 256)), dim3(256), 0, ctx->stream, n, n_pad,
                       x, y, z, w, wscale, (double *)p);
    IMP_CHECK_LAUNCH();
    *out = (const double *)p;
    return IMP_OK;
}

int build_spos(imp_ctx *ctx, int n, const float *x, const float *y, const float *z,
               const float *w, double wscale, const double **out) {
    Scratch S{ctx};
    return build_spos(ctx, S, n, x, y, z, w, wscale, out);
}

// ---- stacks of 2D trajectories (F^H d separable in z) -----------------------
// A stack: M = P K samples where partition p (samples [pK, pK + K)) repeats
// the kx, ky and t of partition 0 and holds one kz (stack-of-spirals and
// stack-of-stars acquisitions).  Then
//   F^H d(r) = sum_c conj(S_c(r)) sum_p e^{+i 2pi kz_p z} G_{c,p}(r),
//   G_{c,p}(r) = sum_k conj(phi) d_{c,p,k} e^{+i (2pi (kx_k x + ky_k y) + w(r) t_k)},
// a 2D NUDFT with C P right-hand sides, so the phase and sin/cos of each
// (sample, voxel) pair serve C P columns instead of C.  W = conj(phi) d is
// [C][P][K], i.e. already [(c P + p)][k].  The kz factor is applied in FP64
// in the reduction.

// smallest i with kz[i] != kz[0] (first initialised to n).  One atomic per
// wave at most, and only when it would lower the current value: atomics on
// one address serialise, and every wave past the first partition finds a change.
__global__ void k_first_change(int n, const float *kz, int *first) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int m = (i < n && kz[i] != kz[0]) ? i : n;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = min(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0 && m < n &&
        m < __hip_atomic_load(first, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMin(first, m);
}

// bad = 1 unless every partition repeats partition 0's kx, ky, t (and phi,
// when given) and holds one kz
__global__ void k_stack_check(int n, int K, const float *kx, const float *ky, const float *kz,
                              const float *t, const float *phi_r, const float *phi_i,
                              int *bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = i % K;
    bool same = kx[i] == kx[j] && ky[i] == ky[j] && t[i] == t[j] && kz[i] == kz[i - j];
    if (phi_r) same = same && phi_r[i] == phi_r[j] && phi_i[i] == phi_i[j];
    // one atomic per wave that holds a mismatch
    const unsigned long long miss = __ballot(!same);
    if (miss && (threadIdx.x & 63) == (unsigned)(__ffsll(miss) - 1)) atomicOr(bad, 1);
}

// Partials of the group launches [group][chunk][64][n] (32 rhs per group,
// rhs q = c P + p) -> out = sum_p e^{+i 2pi kz_p z} sum_c conj(S_c) G_{c,p}, in FP64.
__global__ void k_reduce_stack(const float *partial, int nchunks, int ncoils, int P, int K,
                               int n, const unsigned int *maxbits, const float *kz,
                               const float *z, const float *s_r, const float *s_i, float *out_r,
                               float *out_i) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const double unscale = ldexp(1.0, -w_shift(*maxbits));
    const double za = z[a];
    double orr = 0.0, oi = 0.0;
    for (int p = 0; p < P; ++p) {
        double sp, cp;
        sincospi(2.0 * ((double)kz[(size_t)p * K] * za), &sp, &cp);
        double gr = 0.0, gi = 0.0;
        for (int c = 0; c < ncoils; ++c) {
            const int q = c * P + p, g = q / (2 * MF_COLS), col = q % (2 * MF_COLS);
            const int tt = col / MF_COLS, cc = col % MF_COLS;
            double vr = 0.0, vi = 0.0;
            for (int ch = 0; ch < nchunks; ++ch) {
                const size_t base = ((size_t)g * nchunks + ch) * 64 + tt * 32;
                vr += partial[(base + cc) * n + a];
                vi += partial[(base + MF_COLS + cc) * n + a];
            }
            double sr = 1.0, si = 0.0;
            if (s_r) {
                sr = s_r[(size_t)c * n + a];
                si = s_i[(size_t)c * n + a];
            }
            gr += sr * vr + si * vi;
            gi += sr * vi - si * vr;
        }
        orr += cp * gr - sp * gi;
        oi += cp * gi + sp * gr;
    }
    out_r[a] = (float)(orr * unscale);
    out_i[a] = (float)(oi * unscale);
}

// Partition length K when the samples form a stack (phi, if given, repeats
// too), else 0.
int find_stack(imp_ctx *ctx, int num_k, const float *kx, const float *ky, const float *kz,
               const float *t, const float *phi_r, const float *phi_i, int *K_out) {
    *K_out = 0;
    if (!kz || num_k < 2 * MF_K) return IMP_OK;
    const hipStream_t st = ctx->stream;
    void *w = nullptr;
    IMP_TRY(workspace(ctx, WS_STACK, 2 * sizeof(int), &w));
    int *dev = (int *)w;
    int h[2] = {num_k, 0};
    IMP_HIP(hipMemcpyAsync(dev, h, sizeof(h), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_first_change, dim3(ceil_div(num_k, 256)), dim3(256), 0, st, num_k, kz,
                       dev);
    IMP_CHECK_LAUNCH();
    IMP_HIP(hipMemcpyAsync(h, dev, sizeof(int), hipMemcpyDeviceToHost, st));
    IMP_HIP(hipStreamSynchronize(st));
    const int K = h[0];
    if (K >= num_k || K < MF_K || num_k % K) return IMP_OK;
    hipLaunchKernelGGL(k_stack_check, dim3(ceil_div(num_k, 256)), dim3(256), 0, st, num_k, K,
                       kx, ky, kz, t, phi_r, phi_i, dev + 1);
    IMP_CHECK_LAUNCH();
    IMP_HIP(hipMemcpyAsync(&h[1], dev + 1, sizeof(int), hipMemcpyDeviceToHost, st));
    IMP_HIP(hipStreamSynchronize(st));
    if (!h[1]) *K_out = K;
    return IMP_OK;
}

// Partition length for the separable F^H d, or 0 when the samples are not a
// stack, the separable form would not save at least half of the phase
// evaluations, or its partials would exceed STACK_PART_CAP.
constexpr size_t STACK_PART_CAP = (size_t)8 << 30;

int detect_stack(imp_ctx *ctx, int num_k, int ncoils, int num_x, const float *kx,
                 const float *ky, const float *kz, const float *t, int *K_out) {
    *K_out = 0;
    int K = 0;
    IMP_TRY(find_stack(ctx, num_k, kx, ky, kz, t, nullptr, nullptr, &K));
    if (K == 0) return IMP_OK;
    const int P = num_k / K;
    const long groups = ceil_div((long)ncoils * P, 2 * MF_COLS);
    if (2 * groups > P) return IMP_OK;
    const long wgs = ceil_div(num_x, MF_WAVE_PTS * (NU_THREADS / 64));
    const int chunks = pick_chunks(ctx, wgs, ceil_div(K, MF_K), 4 * MF_FOLD);
    if ((size_t)groups * chunks * 64 * num_x * sizeof(float) > STACK_PART_CAP) return IMP_OK;
    *K_out = K;
    return IMP_OK;
}

// F^H d of a stack (W = conj(phi) d built, lanes built with the field map).
int run_stack_adjoint(imp_ctx *ctx, int num_k, int K, int ncoils, const float4 *lanes, int n,
                      const float *kx, const float *ky, const float *kz, const float *t,
                      const float *z, const float2 *W, const float *s_r, const float *s_i,
                      float *out_r, float *out_i) {
    const hipStream_t st = ctx->stream;
    const int P = num_k / K;
    const int nrhs = ncoils * P;
    const int groups = ceil_div(nrhs, 2 * MF_COLS);
    constexpr int tiles = 2;
    const double *spos = nullptr;
    IMP_TRY(build_spos(ctx, K, kx, ky, nullptr, t, 1.0, &spos));
    void *mb = nullptr, *bp = nullptr, *part = nullptr;
    IMP_TRY(workspace(ctx, WS_MAXBITS, 256, &mb));
    unsigned int *maxbits = (unsigned int *)mb;
    IMP_HIP(hipMemsetAsync(maxbits, 0, sizeof(unsigned int), st));
    hipLaunchKernelGGL(k_absmax, dim3(512), dim3(256), 0, st, W, (size_t)nrhs * K, maxbits);
    IMP_CHECK_LAUNCH();
    const int n_steps = ceil_div(K, MF_K);
    const size_t nb = (size_t)(n_steps + MF_BPAD) * tiles * 64;
    IMP_TRY(workspace(ctx, WS_BPACK, nb * sizeof(BPair), &bp));
    const long wgs = ceil_div(n, MF_WAVE_PTS * (NU_THREADS / 64));
    const int chunks = pick_chunks(ctx, wgs, n_steps, 4 * MF_FOLD);
    const int spc = ceil_div(n_steps, chunks);
    const size_t per_group = (size_t)chunks * tiles * 32 * n;
    IMP_TRY(workspace(ctx, WS_PARTIAL, groups * per_group * sizeof(float), &part));
    for (int g = 0; g < groups; ++g) {
        const int nrhs_g = std::min(2 * MF_COLS, nrhs - g * 2 * MF_COLS);
        hipLaunchKernelGGL(k_pack_mfma, dim3(ceil_div((long)nb, 256)), dim3(256), 0, st, K,
                           nrhs_g, tiles, n_steps + MF_BPAD, 1.f, W + (size_t)g * 2 * MF_COLS * K,
                           maxbits, (BPair *)bp);
        IMP_CHECK_LAUNCH();
        const int rec = prof_begin(ctx, "k_nudft_stack");
        hipLaunchKernelGGL(pick_mfma(tiles, false), dim3((unsigned)wgs, chunks), dim3(NU_THREADS),
                           0, st, lanes, n, spos, (const BPair *)bp, n_steps, spc,
                           (float *)part + g * per_group, 1.0f);
        prof_end(ctx, rec);
        IMP_CHECK_LAUNCH();
    }
    const int rec = prof_begin(ctx, "k_reduce_stack");
    hipLaunchKernelGGL(k_reduce_stack, dim3(ceil_div(n, 256)), dim3(256), 0, st,
                       (const float *)part, chunks, ncoils, P, K, n, maxbits, kz, z, s_r, s_i,
                       out_r, out_i);
    prof_end(ctx, rec);
    IMP_CHECK_LAUNCH();
    return IMP_OK;
}

}  // namespace

// ---- operator entry points (device pointers) --------------------------------

namespace {

bool adjoint_args_ok(imp_ctx *ctx, int num_k, int num_x, int ncoils, const float *kx,
                     const float *ky, const float *t, const float *x, const float *y,
                     const float *phi_r, const float *phi_i, const float *d_r, const float *d_i,
                     const float *s_r, const float *s_i, const float *out_r, const float *out_i) {
    return ctx && num_k > 0 && num_x > 0 && ncoils > 0 && kx && ky && t && x && y && d_r &&
           d_i && out_r && out_i && !(s_r && !s_i) && !(phi_r && !phi_i);
}

// F^H d with the scratch of S; stacked trajectories take the separable form
// only when allow_stack (its detection synchronises with the host).
int adjoint_impl(imp_ctx *ctx, Scratch &S, bool allow_stack, int num_k, int num_x, int ncoils,
                 const float *kx, const float *ky, const float *kz, const float *t,
                 const float *x, const float *y, const float *z, const float *fm,
                 const float *phi_r, const float *phi_i, const float *d_r, const float *d_i,
                 const float *s_r, const float *s_i, float *out_r, float *out_i) {
    const bool d3 = kz && z;
    int K = 0;
    if (d3 && allow_stack) IMP_TRY(detect_stack(ctx, num_k, ncoils, num_x, kx, ky, kz, t, &K));
    const double *spos = nullptr;
    if (K == 0) IMP_TRY(build_spos(ctx, S, num_k, kx, ky, d3 ? kz : nullptr, t, 1.0, &spos));
    void *lanes = nullptr, *W = nullptr;
    IMP_TRY(S.get(WS_LANES, (size_t)num_x * sizeof(float4), &lanes));
    hipLaunchKernelGGL(k_lanes_from_arrays, dim3(ceil_div(num_x, 256)), dim3(256), 0,
                       ctx->stream, num_x, x, y, d3 ? z : nullptr, fm, 1.0 / (2.0 * M_PI),
                       (float4 *)lanes);
    IMP_CHECK_LAUNCH();
    IMP_TRY(S.get(WS_W, (size_t)num_k * ncoils * sizeof(float2), &W));
    hipLaunchKernelGGL(k_w_adjoint, dim3(ceil_div((long)num_k * ncoils, 256)), dim3(256), 0,
                       ctx->stream, num_k, ncoils, phi_r, phi_i, d_r, d_i, (float2 *)W);
    IMP_CHECK_LAUNCH();
    if (K > 0)
        return run_stack_adjoint(ctx, num_k, K, ncoils, (const float4 *)lanes, num_x, kx, ky,
                                 kz, t, z, (const float2 *)W, s_r, s_i, out_r, out_i);
    Job J{(const float4 *)lanes, num_x, spos, num_k, (const float2 *)W, ncoils, d3, +1.f,
          Out::Combine, s_r, s_i, out_r, out_i};
    return run_job(ctx, J, S);
}

// Bytes adjoint_impl carves from a fixed region without the stacked form,
// with one chunk of partials (more chunks are used when they fit).
size_t adjoint_fixed_bytes(int num_k, int num_x, int ncoils, bool d3) {
    auto u = Scratch::unit;
    size_t b = u(spos_bytes(num_k, d3)) + u((size_t)num_x * sizeof(float4)) +
               u((size_t)num_k * ncoils * sizeof(float2));
    if (ncoils >= MFMA_MIN_RHS) {
        const int tiles = ceil_div(ncoils, MF_COLS);
        const size_t nb = (size_t)(ceil_div(num_k, MF_K) + MF_BPAD) * tiles * 64;
        b += u(256) + u(nb * sizeof(BPair)) + u((size_t)tiles * 32 * num_x * sizeof(float));
    } else {
        const int G = ncoils <= 1 ? 1 : 2;
        b += u((size_t)num_k * 2 * G * sizeof(float)) + u((size_t)num_x * sizeof(float2));
    }
    return b;
}

}  // namespace

int nudft_adjoint(imp_ctx *ctx, int num_k, int num_x, int ncoils, const float *kx,
                  const float *ky, const float *kz, const float *t, const float *x,
                  const float *y, const float *z, const float *fm,
                  const float *phi_r, const float *phi_i, const float *d_r,
                  const float *d_i, const float *s_r, const float *s_i,
                  float *out_r, float *out_i) {
    if (!adjoint_args_ok(ctx, num_k, num_x, ncoils, kx, ky, t, x, y, phi_r, phi_i, d_r, d_i,
                         s_r, s_i, out_r, out_i))
        return fail(IMP_EINVAL, "nudft_adjoint: bad arguments");
    Scratch S{ctx};
    return adjoint_impl(ctx, S, true, num_k, num_x, ncoils, kx, ky, kz, t, x, y, z, fm, phi_r,
                        phi_i, d_r, d_i, s_r, s_i, out_r, out_i);
}

// ---- the allocation-free F^H d ------------------------------------------------
// Every buffer of imp_nudft_adjoint_noalloc lives in this fixed device array
// (one copy per GPU, held until the process ends).  Host code never passes
// `scratch` itself to a kernel: it uses the device address hipGetSymbolAddress
// returns for the current device.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mx;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// scratch's address on ctx's device, looked up once per device with that
// device current
int noalloc_scratch(imp_ctx *ctx, void **out) {
    std::lock_guard<std::mutex> lk(g_scratch_mx);
    if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES)
        return fail(IMP_EINVAL, "imp_nudft_adjoint_noalloc: device %d out of range", ctx->device);
    if (!g_scratch_ptr[ctx->device]) {
        int cur = -1;
        IMP_HIP(hipGetDevice(&cur));
        if (cur != ctx->device) IMP_HIP(hipSetDevice(ctx->device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != ctx->device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return fail(IMP_EHIP, "imp_nudft_adjoint_noalloc: hipGetSymbolAddress: %s",
                        hipGetErrorString(e));
        g_scratch_ptr[ctx->device] = p;
    }
    *out = g_scratch_ptr[ctx->device];
    return IMP_OK;
}
}  // namespace

int nudft_adjoint_noalloc(imp_ctx *ctx, int num_k, int num_x, int ncoils, const float *kx,
                          const float *ky, const float *kz, const float *t, const float *x,
                          const float *y, const float *z, const float *fm,
                          const float *phi_r, const float *phi_i, const float *d_r,
                          const float *d_i, const float *s_r, const float *s_i,
                          float *out_r, float *out_i) {
    if (!adjoint_args_ok(ctx, num_k, num_x, ncoils, kx, ky, t, x, y, phi_r, phi_i, d_r, d_i,
                         s_r, s_i, out_r, out_i))
        return fail(IMP_EINVAL, "imp_nudft_adjoint_noalloc: bad arguments");
    if (ncoils > 2 * MF_COLS)
        return fail(IMP_EUNSUPPORTED, "imp_nudft_adjoint_noalloc: %d coils (at most 32)", ncoils);
    const size_t need = adjoint_fixed_bytes(num_k, num_x, ncoils, kz && z);
    if (need > NOALLOC_SCRATCH_BYTES)
        return fail(IMP_EUNSUPPORTED,
                    "imp_nudft_adjoint_noalloc: %d samples x %d voxels x %d coils need %zu bytes "
                    "of scratch; the fixed region holds %zu (use imp_nudft_adjoint)",
                    num_k, num_x, ncoils, need, NOALLOC_SCRATCH_BYTES);
    void *base = nullptr;
    IMP_TRY(noalloc_scratch(ctx, &base));
    Scratch S{ctx, (char *)base, NOALLOC_SCRATCH_BYTES, 0};
    return adjoint_impl(ctx, S, false, num_k, num_x, ncoils, kx, ky, kz, t, x, y, z, fm, phi_r,
                        phi_i, d_r, d_i, s_r, s_i, out_r, out_i);
}

// F^H d of several images (different sizes, trajectories, coil maps) with
// one persistent MFMA launch over all of them.  Every job needs 3..32 coils
// and the same dimensionality.  chunk_scale multiplies the per-job
// workgroup target (the autotuned knob; 0 = default 8 per CU).
int nudft_adjoint_batch(imp_ctx *ctx, int njobs, const imp_nudft_job *jobs, int chunk_scale) {
    if (!ctx || njobs <= 0 || !jobs) return fail(IMP_EINVAL, "nudft_adjoint_batch: bad arguments");
    const bool d3 = jobs[0].kz && jobs[0].z;
    int tiles = 0;
    for (int j = 0; j < njobs; ++j) {
        const imp_nudft_job &b = jobs[j];
        if (b.num_k <= 0 || b.num_x <= 0 || !b.kx || !b.ky || !b.t || !b.x || !b.y || !b.d_r ||
            !b.d_i || !b.out_r || !b.out_i || (b.s_r && !b.s_i) || (b.phi_r && !b.phi_i))
            return fail(IMP_EINVAL, "nudft_adjoint_batch: bad job %d", j);
        if (b.ncoils < MFMA_MIN_RHS || b.ncoils > 2 * MF_COLS)
            return fail(IMP_EUNSUPPORTED, "nudft_adjoint_batch: job %d has %d coils (3..32)", j,
                        b.ncoils);
        if ((b.kz && b.z) != d3)
            return fail(IMP_EINVAL, "nudft_adjoint_batch: mixed 2D and 3D jobs");
        const int t = ceil_div(b.ncoils, MF_COLS);
        if (j == 0) tiles = t;
        if (t != tiles)
            return fail(IMP_EUNSUPPORTED, "nudft_adjoint_batch: jobs need the same coil tile count");
    }
    const long target = (long)(chunk_scale > 0 ? chunk_scale : 8) * ctx->num_cus;
    // carve one workspace: per job spos, lanes, W, maxbits, bpack, partials
    struct Lay {
        size_t spos, lanes, w, mb, bp, part;
        int n_steps, chunks, spc, wgs_x;
    };
    std::vector<Lay> lay(njobs);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    for (int j = 0; j < njobs; ++j) {
        const imp_nudft_job &b = jobs[j];
        Lay &L = lay[j];
        L.n_steps = ceil_div(b.num_k, MF_K);
        L.wgs_x = ceil_div(b.num_x, MF_WAVE_PTS * (NU_THREADS / 64));
        L.chunks = 1;
        while ((long)L.wgs_x * L.chunks < target && L.n_steps / (L.chunks * 2) >= 4 * MF_FOLD)
            L.chunks *= 2;
        L.spc = ceil_div(L.n_steps, L.chunks);
        L.spos = take((size_t)L.n_steps * pos_stride(d3) * sizeof(double));
        L.lanes = take((size_t)b.num_x * sizeof(float4));
        L.w = take((size_t)b.num_k * b.ncoils * sizeof(float2));
        L.mb = take(sizeof(unsigned int));
        L.bp = take((size_t)(L.n_steps + MF_BPAD) * tiles * 64 * sizeof(BPair));
        L.part = take((size_t)L.chunks * tiles * 32 * b.num_x * sizeof(float));
    }
    const size_t table = take((size_t)njobs * sizeof(MfmaJob));
    void *wsp = nullptr;
    IMP_TRY(workspace(ctx, WS_BATCH, off, &wsp));
    char *base = (char *)wsp;
    const hipStream_t st = ctx->stream;
    std::vector<MfmaJob> tab(njobs);
    int wg = 0;
    for (int j = 0; j < njobs; ++j) {
        const imp_nudft_job &b = jobs[j];
        const Lay &L = lay[j];
        double *spos = (double *)(base + L.spos);
        float4 *lanes = (float4 *)(base + L.lanes);
        float2 *W = (float2 *)(base + L.w);
        unsigned int *mb = (unsigned int *)(base + L.mb);
        BPair *bp = (BPair *)(base + L.bp);
        const int n_pad = L.n_steps * MF_K;
        hipLaunchKernelGGL(k_spos, dim3(ceil_div(n_pad, 256)), dim3(256), 0, st, b.num_k, n_pad,
                           b.kx, b.ky, d3 ? b.kz : nullptr, b.t, 1.0, spos);
        hipLaunchKernelGGL(k_lanes_from_arrays, dim3(ceil_div(b.num_x, 256)), dim3(256), 0, st,
                           b.num_x, b.x, b.y, d3 ? b.z : nullptr, b.fm, 1.0 / (2.0 * M_PI), lanes);
        hipLaunchKernelGGL(k_w_adjoint, dim3(ceil_div((long)b.num_k * b.ncoils, 256)), dim3(256),
                           0, st, b.num_k, b.ncoils, b.phi_r, b.phi_i, b.d_r, b.d_i, W);
        IMP_HIP(hipMemsetAsync(mb, 0, sizeof(unsigned int), st));
        hipLaunchKernelGGL(k_absmax, dim3(512), dim3(256), 0, st, (const float2 *)W,
                           (size_t)b.ncoils * b.num_k, mb);
        const size_t nb = (size_t)(L.n_steps + MF_BPAD) * tiles * 64;
        hipLaunchKernelGGL(k_pack_mfma, dim3(ceil_div((long)nb, 256)), dim3(256), 0, st, b.num_k,
                           b.ncoils, tiles, L.n_steps + MF_BPAD, +1.f, (const float2 *)W,
                           (const unsigned int *)mb, bp);
        IMP_CHECK_LAUNCH();
        MfmaJob &J = tab[j];
        J.lanes = lanes;
        J.spos = spos;
        J.bpack = bp;
        J.partial = (float *)(base + L.part);
        J.n_lane = b.num_x;
        J.n_steps = L.n_steps;
        J.steps_per_chunk = L.spc;
        J.wg_begin = wg;
        J.wgs_x = L.wgs_x;
        wg += L.wgs_x * L.chunks;
    }
    MfmaJob *dtab = (MfmaJob *)(base + table);
    IMP_HIP(hipMemcpyAsync(dtab, tab.data(), (size_t)njobs * sizeof(MfmaJob),
                           hipMemcpyHostToDevice, st));
    using BFn = void (*)(const MfmaJob *, int, float);
    BFn fn = tiles == 1 ? (d3 ? (BFn)k_nudft_mfma_batch<1, true> : (BFn)k_nudft_mfma_batch<1, false>)
                        : (d3 ? (BFn)k_nudft_mfma_batch<2, true> : (BFn)k_nudft_mfma_batch<2, false>);
    const int rec = prof_begin(ctx, "k_nudft_mfma_batch");
    hipLaunchKernelGGL(fn, dim3(wg), dim3(NU_THREADS), 0, st, (const MfmaJob *)dtab, njobs, 1.0f);
    prof_end(ctx, rec);
    IMP_CHECK_LAUNCH();
    for (int j = 0; j < njobs; ++j) {
        const imp_nudft_job &b = jobs[j];
        const Lay &L = lay[j];
        hipLaunchKernelGGL(k_reduce_mfma, dim3(ceil_div(b.num_x, 256)), dim3(256), 0, st,
                           (const float *)(base + L.part), L.chunks, tiles, b.ncoils, b.num_x, 0,
                           (const unsigned int *)(base + L.mb), b.s_r, b.s_i, b.out_r, b.out_i);
    }
    IMP_CHECK_LAUNCH();
    // the host job table must outlive the copy
    IMP_HIP(hipStreamSynchronize(st));
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
    // a stack: a 2D forward NUDFT of one partition with the C P coil-partition
    // pairs as right-hand sides, the kz factor folded into their weights
    int K = 0;
    if (d3) {
        IMP_TRY(find_stack(ctx, num_k, kx, ky, kz, t, nullptr, nullptr, &K));
        if (K > 0) {
            const int P = num_k / K;
            const long groups = ceil_div((long)ncoils * P, 2 * MF_COLS);
            if (2 * groups > P || (size_t)ncoils * P * num_x > ((size_t)1 << 30)) K = 0;
        }
    }
    if (K > 0) {
        const int P = num_k / K, nrhs = ncoils * P;
        const double *spos = nullptr;
        IMP_TRY(build_spos(ctx, num_x, x, y, nullptr, fm, 1.0 / (2.0 * M_PI), &spos));
        void *lanes = nullptr, *W = nullptr;
        IMP_TRY(workspace(ctx, WS_LANES, (size_t)K * sizeof(float4), &lanes));
        hipLaunchKernelGGL(k_lanes_from_arrays, dim3(ceil_div(K, 256)), dim3(256), 0,
                           ctx->stream, K, kx, ky, nullptr, t, 1.0, (float4 *)lanes);
        IMP_CHECK_LAUNCH();
        const size_t nw = (size_t)nrhs * num_x;
        IMP_TRY(workspace(ctx, WS_W, nw * sizeof(float2), &W));
        hipLaunchKernelGGL(k_w_forward_stack, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0,
                           ctx->stream, num_x, ncoils, P, K, kz, z, rho_r, rho_i, s_r, s_i,
                           (float2 *)W);
        IMP_CHECK_LAUNCH();
        for (int q0 = 0; q0 < nrhs; q0 += 2 * MF_COLS) {
            const int nq = std::min(2 * MF_COLS, nrhs - q0);
            Job J{(const float4 *)lanes, K, spos, num_x, (const float2 *)W + (size_t)q0 * num_x,
                  nq, false, -1.f, Out::PerRhs, nullptr, nullptr, d_r + (size_t)q0 * K,
                  d_i + (size_t)q0 * K, "k_nudft_stack"};
            IMP_TRY(run_job(ctx, J));
        }
        return IMP_OK;
    }
    const double *spos = nullptr;
    IMP_TRY(build_spos(ctx, num_x, x, y, d3 ? z : nullptr, fm, 1.0 / (2.0 * M_PI), &spos));
    void *lanes = nullptr, *W = nullptr;
    IMP_TRY(workspace(ctx, WS_LANES, (size_t)num_k * sizeof(float4), &lanes));
    hipLaunchKernelGGL(k_lanes_from_arrays, dim3(ceil_div(num_k, 256)), dim3(256), 0,
                       ctx->stream, num_k, kx, ky, d3 ? kz : nullptr, t, 1.0, (float4 *)lanes);
    IMP_CHECK_LAUNCH();
    IMP_TRY(workspace(ctx, WS_W, (size_t)num_x * ncoils * sizeof(float2), &W));
    hipLaunchKernelGGL(k_w_forward, dim3(ceil_div((long)num_x * ncoils, 256)), dim3(256), 0,
                       ctx->stream, num_x, ncoils, rho_r, rho_i, s_r, s_i, (float2 *)W);
    IMP_CHECK_LAUNCH();
    Job J{(const float4 *)lanes, num_k, spos, num_x, (const float2 *)W, ncoils, d3, -1.f,
          Out::PerRhs, nullptr, nullptr, d_r, d_i};
    return run_job(ctx, J);
}

namespace {

// ---- Q on the standard grid as a separable GEMM ------------------------------
//
// On the standard grid e^{i 2pi k.r} = ex(k, x) ey(k, y) ez(k, z), so
//     Q_l[y][x][z] = sum_k (a_l(t_k) |phi_k|^2 ey(k, y) ez(k, z)) ex(k, x),
// a complex GEMM with rows (y, z), columns x and the samples as K.  Its
// operands come from per-sample phasor tables: one FP64 sincos per sample
// and grid line instead of one per sample and grid point.  a_l(t) is
// nonzero only for (t - t0)/tau in (l - 1, l + 1), so a stable counting
// sort puts the samples in buckets b = floor((t - t0)/tau) + 1 (each padded
// to a whole step of 8 samples) and segment l runs over buckets l and l + 1
// only: about 2/L of the samples instead of all of them.
// Rows y < 0 (iy < Ny) are computed.  Rows y > 0 are their mirrors
// Q(-r) = conj(Q(r)) (the weights are real), except where the mirror leaves
// the grid (x = -Nx, z = -Nz); those points and the row y = 0 go through the
// direct per-point kernel.
// GEMM layout: one wave owns 64 rows, two 32-row M tiles whose A operands
// are built in registers from the ey/ez tables and the wave-uniform weights
// of its segment, split into fp16 hi/lo and swapped into [Re | Im] K halves
// as in k_nudft_mfma.  The wave also owns QS_TILES 16-column N tiles whose
// B operands (ex, split, arranged per lane) the WM waves of a workgroup share
// through LDS.  A workgroup covers one K block of one segment; the K-block
// partials are reduced in FP64.

constexpr int QS_TILES = 4;     // N tiles (16 complex columns each) per wave
constexpr int QS_BLK = 1024;    // samples per counting-sort block
constexpr int QS_MAXBK = 33;    // buckets: nseg + 1 <= 33
constexpr int QS_BSHIFT = 10;   // ex is stored scaled by 2^10 (keeps lo parts normal)
constexpr size_t QS_PART_CAP = (size_t)4 << 30;   // K-block partials, bytes
constexpr size_t QS_TAB_CAP = (size_t)64 << 30;   // phasor tables; above: direct kernel
constexpr size_t QS_KEEP = (size_t)1 << 30;       // workspaces above this are freed after use
constexpr int QS_FALLBACK = 1;                    // toeplitz_q_sep: use the direct kernel

struct QsJob {
    int l;       // segment
    int s0, s1;  // steps [s0, s1) of the bucketed sample order
    int pad;
};

// bucket of a sample; -1 when a_l(t) = 0 for every l
__device__ __forceinline__ int q_bucket(float t, float tau, float t0, int nseg) {
    const double a = ((double)t - (double)t0) / (double)tau;
    if (!(a > -1.0) || !(a < (double)nseg)) return -1;
    return (int)floor(a) + 1;
}

__global__ void k_fill_i32(int *p, long n, int v) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// per block of QS_BLK samples: cnt[b][blk]
__global__ __launch_bounds__(256) void k_qb_count(int n, const float *t, float tau, float t0,
                                                  int nseg, int nblk, int *cnt) {
    __shared__ int h[QS_MAXBK];
    const int nbk = nseg + 1;
    for (int i = threadIdx.x; i < nbk; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < QS_BLK; i += blockDim.x) {
        const long k = (long)blockIdx.x * QS_BLK + i;
        if (k < n) {
            const int b = q_bucket(t[k], tau, t0, nseg);
            if (b >= 0) atomicAdd(&h[b], 1);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nbk; i += blockDim.x) cnt[(size_t)i * nblk + blockIdx.x] = h[i];
}

// counts -> positions (bucket-major; every bucket padded to a whole step):
// cnt[b][blk] becomes the position of block blk's first sample of bucket b;
// bstart[b] = first position of bucket b, bstart[nbk] = padded total.
// One workgroup of 1024 threads scans each bucket's block counts.
__global__ __launch_bounds__(1024) void k_qb_scan(int nblk, int nbk, int *cnt, int *bstart) {
    __shared__ int wsum[16];
    __shared__ int carry;
    __shared__ int start[QS_MAXBK + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) start[0] = 0;
    for (int b = 0; b < nbk; ++b) {
        if (tid == 0) carry = 0;
        __syncthreads();
        for (int c0 = 0; c0 < nblk; c0 += 1024) {
            const int k = c0 + tid;
            const int v = k < nblk ? cnt[(size_t)b * nblk + k] : 0;
            int x = v;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int y = __shfl_up(x, d, 64);
                if (lane >= d) x += y;
            }
            if (lane == 63) wsum[wave] = x;
            __syncthreads();
            if (wave == 0) {
                int w = lane < 16 ? wsum[lane] : 0;
#pragma unroll
                for (int d = 1; d < 16; d <<= 1) {
                    const int y = __shfl_up(w, d, 64);
                    if (lane >= d) w += y;
                }
                if (lane < 16) wsum[lane] = w;
            }
            __syncthreads();
            if (k < nblk)
                cnt[(size_t)b * nblk + k] = start[b] + carry + (wave > 0 ? wsum[wave - 1] : 0) + x - v;
            __syncthreads();
            if (tid == 0) carry += wsum[15];
            __syncthreads();
        }
        if (tid == 0) start[b + 1] = start[b] + (carry + MF_K - 1) / MF_K * MF_K;
        __syncthreads();
    }
    if (tid <= nbk) bstart[tid] = start[tid];
}

// stable scatter: perm[position] = sample index (padding keeps -1)
__global__ __launch_bounds__(256) void k_qb_scatter(int n, const float *t, float tau, float t0,
                                                    int nseg, int nblk, const int *off,
                                                    int *perm) {
    __shared__ int base[QS_MAXBK];
    __shared__ int wcnt[4][QS_MAXBK];
    const int nbk = nseg + 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < nbk; i += 256) base[i] = off[(size_t)i * nblk + blockIdx.x];
    __syncthreads();
    for (int it = 0; it < QS_BLK / 256; ++it) {
        const long k = (long)blockIdx.x * QS_BLK + it * 256 + threadIdx.x;
        const int b = k < n ? q_bucket(t[k], tau, t0, nseg) : -1;
        int rank = 0;
        for (int c = 0; c < nbk; ++c) {
            const unsigned long long m = __ballot(b == c);
            if (b == c) rank = __popcll(m & ((1ull << lane) - 1ull));
            if (lane == 0) wcnt[wave][c] = __popcll(m);
        }
        __syncthreads();
        if (b >= 0) {
            int p = base[b] + rank;
            for (int w = 0; w < wave; ++w) p += wcnt[w][b];
            perm[p] = (int)k;
        }
        __syncthreads();
        if (threadIdx.x < nbk)
            base[threadIdx.x] += wcnt[0][threadIdx.x] + wcnt[1][threadIdx.x] +
                                 wcnt[2][threadIdx.x] + wcnt[3][threadIdx.x];
        __syncthreads();
    }
}

// wq[l][p] = a_l(t) |phi|^2 2^shift of the sample at position p (0 on padding)
__global__ void k_qt_w(int kp, int nseg, int num_k, const int *perm, const float2 *W,
                       const unsigned int *maxbits, float *wq) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)kp * nseg) return;
    const int l = (int)(i / kp), p = (int)(i % kp);
    const int k = perm[p];
    const float scale = ldexpf(1.f, w_shift(*maxbits));
    wq[i] = k >= 0 ? W[(size_t)l * num_k + k].x * scale : 0.f;
}

// revolutions kc (i - Nc)/Nc in FP64, reduced to [-1/2, 1/2]
__device__ __forceinline__ double q_rev(float kc, int i, int Nc) {
    const double u = (double)kc * (double)(i - Nc) / (double)Nc;
    return u - rint(u);
}

// tab[step][line][j] = e^{i 2pi kc (line - Nc)/Nc} of the sample at
// position 8 step + j, lines [0, nl); 0 on padding
__global__ void k_qt_line(int kp, int nl, int Nc, const int *perm, const float *kc,
                          float2 *tab) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)kp * nl) return;
    const int j = (int)(i & (MF_K - 1));
    const long sl = i / MF_K;  // step * nl + line
    const int line = (int)(sl % nl);
    const long step = sl / nl;
    const int k = perm[step * MF_K + j];
    float2 v = make_float2(0.f, 0.f);
    if (k >= 0) {
        double s, c;
        sincospi(2.0 * q_rev(kc[k], line, Nc), &s, &c);
        v = make_float2((float)c, (float)s);
    }
    tab[i] = v;
}

__device__ __forceinline__ void split_h(double v, _Float16 &hi, _Float16 &lo) {
    hi = (_Float16)(float)v;
    lo = (_Float16)(float)(v - (double)(float)hi);
}

// B operands of the Q GEMM, [step][tile][hi, lo][lane] half8, ntp tiles of
// 16 x values.  Lane l: column n = l%32 -> part n/16 of the product (Re, Im)
// and x = 16 tile + n%16; K rows (h = l/32, sample j): h = 0 multiplies
// Re A, h = 1 multiplies Im A, so Re = Ar Br - Ai Bi, Im = Ar Bi + Ai Br.
// A workgroup covers one step and 16 tiles (256 x values, one per thread).
// e^{i 2pi kx (x - Nx)/Nx} = e^{i 2pi kx xl/Nx} e^{i 2pi kx (16 tile - Nx)/Nx}
// (x = 16 tile + xl): the workgroup evaluates both factors of its 8 samples
// in FP64 (one sincos per thread) into LDS, and each thread forms its x by
// one complex FP64 product per sample, then writes its 4 lanes' half8.
__global__ __launch_bounds__(256) void k_qt_bx(int ntp, int C2, int Nx, const int *perm,
                                               const float *kx, half8 *bx) {
    __shared__ double2 f_lo[MF_K][16];  // e^{i 2pi kx xl / Nx}
    __shared__ double2 f_hi[MF_K][16];  // e^{i 2pi kx (16 tile - Nx) / Nx}
    const long step = blockIdx.x;
    const int tb = blockIdx.y * 16;
    const int tid = threadIdx.x;
    {
        const int j = tid >> 5, hi = (tid >> 4) & 1, idx = tid & 15;
        const int k = perm[step * MF_K + j];
        double sn = 0.0, cs = 0.0;
        if (k >= 0) {
            const double kk = (double)kx[k];
            double u = hi ? kk * (double)(16 * (tb + idx) - Nx) / (double)Nx
                          : kk * (double)idx / (double)Nx;
            sincospi(2.0 * (u - rint(u)), &sn, &cs);
        }
        (hi ? f_hi : f_lo)[j][idx] = make_double2(cs, sn);
    }
    __syncthreads();
    const int tile = tb + (tid >> 4), xl = tid & 15;
    if (tile >= ntp) return;
    const bool in = tile * 16 + xl < C2;
    const double sc = (double)(1 << QS_BSHIFT);
    half8 ch, cl, sh, sl, nsh, nsl;
#pragma unroll
    for (int j = 0; j < MF_K; ++j) {
        const double2 a = f_lo[j][xl], b = f_hi[j][tid >> 4];
        const double c = in ? a.x * b.x - a.y * b.y : 0.0;
        const double s = in ? a.x * b.y + a.y * b.x : 0.0;
        _Float16 p, q;
        split_h(c * sc, p, q);
        ch[j] = p;
        cl[j] = q;
        split_h(s * sc, p, q);
        sh[j] = p;
        sl[j] = q;
        nsh[j] = -p;
        nsl[j] = -q;
    }
    half8 *ho = bx + ((size_t)step * ntp + tile) * 2 * 64;
    half8 *lo = ho + 64;
    ho[xl] = ch;       lo[xl] = cl;        // h = 0, Re column: Br
    ho[16 + xl] = sh;  lo[16 + xl] = sl;   // h = 0, Im column: Bi
    ho[32 + xl] = nsh; lo[32 + xl] = nsl;  // h = 1, Re column: -Bi
    ho[48 + xl] = ch;  lo[48 + xl] = cl;   // h = 1, Im column: Br
}

template <int WM>
struct QsLds {
    float cmp[WM][2 * QS_TILES * 16][64];
    half8 b[2][QS_TILES][2][64];
};

// One workgroup = WM waves x 64 rows, QS_TILES N tiles, one K block (job).
// B operands of the next step are staged into the other LDS buffer while
// the current one is used (one barrier per step); global loads run two
// steps ahead of their use.  part[job][Re|Im][R][C2].
template <int WM, bool D3>
__global__ __launch_bounds__(64 * WM) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_q_sep(const float4 *__restrict__ ey, const float4 *__restrict__ ez,
             const float *__restrict__ wq, const half8 *__restrict__ bx,
             const QsJob *__restrict__ jobs, int njobs, int gx, int gy, int R, int Ny, int M3,
             int ntp, int kp, int C2, float *__restrict__ part, float one) {
    constexpr int NA = 2 * QS_TILES;
    constexpr int NT = 64 * WM;
    constexpr int NV = QS_TILES * 2 * 64;  // half8 of B per step and workgroup
    constexpr int NB = (NV + NT - 1) / NT;
    __shared__ QsLds<WM> lds;
    auto &cmp = lds.cmp;
    const int l = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int h = l >> 5;
    // XCD-aware order: workgroups go to the 8 XCDs round-robin, so slot
    // b / 8 of XCD b % 8 takes job 8 * (slot / (gx gy)) + b % 8: every
    // workgroup of a job (all its row and column blocks, which share the A
    // rows and the B tiles) runs on one XCD's L2, column blocks first.
    const int per = gx * gy;
    const int slot = (int)(blockIdx.x >> 3);
    const int job = 8 * (slot / per) + (int)(blockIdx.x & 7);
    if (job >= njobs) return;
    const int bxi = slot % per % gx, byi = slot % per / gx;
    const QsJob J = jobs[job];
    const int row0 = (byi * WM + wave) * 64;
    const int row = row0 + l;
    const bool rv = row < R;
    const int iy = rv ? row / M3 : 0, iz = rv ? row % M3 : 0;
    const int tile0 = bxi * QS_TILES;
    const float *__restrict__ wl = wq + (size_t)J.l * kp;
    auto bsrc = [&](int s) { return bx + ((size_t)s * ntp + tile0) * 2 * 64; };

    floatx16 acc[NA];
    float tot[NA][16];
#pragma unroll
    for (int q = 0; q < NA; ++q) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            acc[q][i] = 0.f;
            tot[q][i] = 0.f;
            cmp[wave][q * 16 + i][l] = 0.f;
        }
    }
    {
        const half8 *src = bsrc(J.s0);
        half8 *dst = &lds.b[0][0][0][0];
        for (int v = threadIdx.x; v < NV; v += NT) dst[v] = src[v];
    }
    // Software pipeline.  Step s issues its MFMAs with A operands built
    // during step s - 1 and, independent of them, builds the A operands of
    // step s + 1 (VALU), so the two interleave.  A data are fetched two steps
    // before they are built on; B operands two steps before their LDS stage
    // (B2) or one (a thread's share of B above 4 vectors, 2 in 3D, or uneven).
    // Register sets alternate by step parity (the loop is unrolled by two),
    // so no register holding a load in flight is copied (a copy waits for
    // the load).  Loads are unconditional, prefetch indices clamped to the
    // last step (a harmless reload at the tail), the B staging unconditional
    // (the buffer written at the last step is never read) and every job has
    // an even number of steps: the step body has no branches, and the
    // wait-count insertion waits only for the loads each use needs.  Lanes
    // of rows past R read row 0's data: their output rows are never written
    // and rows of the GEMM are independent.
    constexpr bool B2 = NB <= (D3 ? 2 : 4) && NV % NT == 0;
    constexpr bool NV_EVEN = NV % NT == 0;
    auto load_b = [&](int s, half8 (&r)[NB]) {
        const half8 *src = bsrc(s);
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int v = threadIdx.x + u * NT;
            if (NV_EVEN || v < NV) r[u] = src[v];
        }
    };
    auto load_a = [&](int s, float4 (&e)[4], float4 (&z)[4]) {
        const float4 *pe = ey + ((size_t)s * Ny + iy) * 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) e[q] = pe[q];
        if constexpr (D3) {
            const float4 *pz = ez + ((size_t)s * M3 + iz) * 4;
#pragma unroll
            for (int q = 0; q < 4; ++q) z[q] = pz[q];
        }
    };
    // A operands of one step: [Re | Im] K halves of M tiles A (rows 0-31)
    // and B (rows 32-63) of this wave, fp16 hi and lo
    struct AOps {
        half8 hA, hB, lA, lB;
    };
    auto build = [&](int s, const float4 (&ae)[4], const float4 (&ze)[4], AOps &o) {
        const float *__restrict__ w8 = wl + (size_t)s * MF_K;
        float arv[MF_K], aiv[MF_K];
#pragma unroll
        for (int j = 0; j < MF_K; ++j) {
            const float4 e4 = ae[j >> 1];
            float pr = (j & 1) ? e4.z : e4.x;
            float pi = (j & 1) ? e4.w : e4.y;
            if constexpr (D3) {
                const float4 z4 = ze[j >> 1];
                const float zr = (j & 1) ? z4.z : z4.x;
                const float zi = (j & 1) ? z4.w : z4.y;
                const float tr = pr * zr - pi * zi;
                pi = pr * zi + pi * zr;
                pr = tr;
            }
            const float w = w8[j];
            float ar = w * pr, ai = w * pi;
            // Opaque to the optimizer, so that the split below sees one
            // FP32 value: a conversion of w * pr re-derived as a fused
            // multiply-convert (v_fma_mixlo_f16, single rounding) next to one
            // of the rounded product gives two different hi parts, and hi + lo
            // then misses w * pr by one fp16 ulp.
            asm("" : "+v"(ar), "+v"(ai));
            arv[j] = ar;
            aiv[j] = ai;
        }
        u32x4 rhu, rlu, ihu, ilu;
        split8(arv, one, rhu, rlu);
        split8(aiv, one, ihu, ilu);
        u32x4 hA, hB, lA, lB;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const auto ph = __builtin_amdgcn_permlane32_swap(rhu[i], ihu[i], false, false);
            hA[i] = ph[0];
            hB[i] = ph[1];
            const auto pl = __builtin_amdgcn_permlane32_swap(rlu[i], ilu[i], false, false);
            lA[i] = pl[0];
            lB[i] = pl[1];
        }
        o.hA = __builtin_bit_cast(half8, hA);
        o.hB = __builtin_bit_cast(half8, hB);
        o.lA = __builtin_bit_cast(half8, lA);
        o.lB = __builtin_bit_cast(half8, lB);
    };

    const int s_last = J.s1 - 1;
    float4 d0[4], z0[4], d1[4], z1[4];
    half8 b0[NB], b1[NB];
    AOps op0, op1;
    load_a(J.s0, d0, z0);
    load_a(min(J.s0 + 1, s_last), d1, z1);
    if constexpr (B2) {
        load_b(min(J.s0 + 1, s_last), b0);
        load_b(min(J.s0 + 2, s_last), b1);
    }
    build(J.s0, d0, z0, op0);
    load_a(min(J.s0 + 2, s_last), d0, z0);
    __syncthreads();

    // step s: MFMAs with `cur`; `nxt` built from dn (data of step s + 1),
    // dn refilled with step s + 3; bq holds B of step s + 1 (B2) and is
    // refilled with step s + 3
    auto step = [&](int s, const AOps &cur, AOps &nxt, float4 (&dn)[4], float4 (&zn)[4],
                    half8 (&bq)[NB]) {
        const int buf = (s - J.s0) & 1;
        if (!B2) load_b(min(s + 1, s_last), bq);
        half8 bh[QS_TILES], bl[QS_TILES];
#pragma unroll
        for (int t = 0; t < QS_TILES; ++t) {
            bh[t] = lds.b[buf][t][0][l];
            bl[t] = lds.b[buf][t][1][l];
        }
        build(min(s + 1, s_last), dn, zn, nxt);
        load_a(min(s + 3, s_last), dn, zn);
#pragma unroll
        for (int t = 0; t < QS_TILES; ++t) {
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(cur.hA, bh[t], acc[t], 0, 0, 0);
            acc[QS_TILES + t] =
                __builtin_amdgcn_mfma_f32_32x32x16_f16(cur.hB, bh[t], acc[QS_TILES + t], 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < QS_TILES; ++t) {
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(cur.hA, bl[t], acc[t], 0, 0, 0);
            acc[QS_TILES + t] =
                __builtin_amdgcn_mfma_f32_32x32x16_f16(cur.hB, bl[t], acc[QS_TILES + t], 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < QS_TILES; ++t) {
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(cur.lA, bh[t], acc[t], 0, 0, 0);
            acc[QS_TILES + t] =
                __builtin_amdgcn_mfma_f32_32x32x16_f16(cur.lB, bh[t], acc[QS_TILES + t], 0, 0, 0);
        }
        {
            half8 *dst = &lds.b[buf ^ 1][0][0][0];
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                const int v = threadIdx.x + u * NT;
                if (NV_EVEN || v < NV) dst[v] = bq[u];
            }
        }
        if constexpr (B2) load_b(min(s + 3, s_last), bq);
        __syncthreads();
    };

    for (int f0 = J.s0; f0 < J.s1; f0 += MF_FOLD) {
        const int f1 = min(J.s1, f0 + MF_FOLD);
        for (int s = f0; s < f1; s += 2) {
            step(s, op0, op1, d1, z1, b0);
            if constexpr (B2) step(s + 1, op1, op0, d0, z0, b1);
            else step(s + 1, op1, op0, d0, z0, b0);
        }
#pragma unroll
        for (int q = 0; q < NA; ++q) {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                float sum, err;
                two_sum(tot[q][i], acc[q][i], sum, err);
                tot[q][i] = sum;
                cmp[wave][q * 16 + i][l] += err;
                acc[q][i] = 0.f;
            }
        }
    }
    // D layout of 32x32 MFMA: lane l, register i -> column l%32,
    // row 8*(i/4) + 4*(l/32) + i%4
    const int n = l & 31;
    const size_t plane = (size_t)R * C2;
    float *__restrict__ out = part + (size_t)job * 2 * plane + (size_t)(n >> 4) * plane;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
#pragma unroll
        for (int t = 0; t < QS_TILES; ++t) {
            const int q = m * QS_TILES + t;
            const int x = (tile0 + t) * 16 + (n & 15);
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int r = row0 + m * 32 + 8 * (i >> 2) + 4 * h + (i & 3);
                if (r < R && x < C2) out[(size_t)r * C2 + x] = tot[q][i] + cmp[wave][q * 16 + i][l];
            }
        }
    }
}

// K-block partials -> Q on the rows iy < Ny and their mirrors; FP64 sums,
// scales undone; cumulative: prefix sums over l.
__global__ void k_qs_reduce(const float *part, const int *jb, int nseg, int R, int C2, int Nx,
                            int Ny, int Nz, int M3, int cumulative, const unsigned int *maxbits,
                            float *q_r, float *q_i) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)R * C2) return;
    const int x = (int)(i % C2), r = (int)(i / C2);
    const int iy = r / M3, iz = r % M3;
    const long nq = 2L * Ny * C2 * M3;
    const long g = ((long)iy * C2 + x) * M3 + iz;
    const bool mir = iy > 0 && x > 0 && (M3 == 1 || iz > 0);
    const long gm = mir ? ((long)(2 * Ny - iy) * C2 + (2 * Nx - x)) * M3 + (M3 > 1 ? 2 * Nz - iz : 0)
                        : 0;
    const double unscale = ldexp(1.0, -w_shift(*maxbits) - QS_BSHIFT);
    const size_t plane = (size_t)R * C2;
    double cr = 0.0, ci = 0.0;
    for (int l = 0; l < nseg; ++l) {
        double sr = 0.0, si = 0.0;
        for (int j = jb[l]; j < jb[l + 1]; ++j) {
            sr += part[(size_t)j * 2 * plane + i];
            si += part[((size_t)j * 2 + 1) * plane + i];
        }
        sr *= unscale;
        si *= unscale;
        if (cumulative) {
            cr += sr;
            ci += si;
            sr = cr;
            si = ci;
        }
        q_r[(size_t)l * nq + g] = (float)sr;
        q_i[(size_t)l * nq + g] = (float)si;
        if (mir) {
            q_r[(size_t)l * nq + gm] = (float)sr;
            q_i[(size_t)l * nq + gm] = (float)-si;
        }
    }
}

// Points the GEMM path evaluates directly: the row y = 0 (iy = Ny), then the
// points of rows y > 0 whose mirror leaves the grid (qhalf_edge order).
__device__ __forceinline__ void qedge_point(int j, int Nx, int Ny, int Nz, int ne_row, int &iy,
                                            int &ix, int &iz) {
    const int M2 = 2 * Nx, M3 = Nz > 1 ? 2 * Nz : 1;
    if (j < M2 * M3) {
        iy = Ny;
        ix = j / M3;
        iz = j % M3;
    } else {
        qhalf_edge(j - M2 * M3, Ny, M2, M3, ne_row, iy, ix, iz);
    }
}

__global__ void k_lanes_qedge(int Nx, int Ny, int Nz, int ne_row, int n_edge, float4 *lanes) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_edge) return;
    int iy, ix, iz;
    qedge_point(j, Nx, Ny, Nz, ne_row, iy, ix, iz);
    lanes[j] = make_float4((float)(ix - Nx) / (float)Nx, (float)(iy - Ny) / (float)Ny,
                           Nz > 1 ? (float)(iz - Nz) / (float)Nz : 0.f, 0.f);
}

__global__ void k_q_edge_fill(int Nx, int Ny, int Nz, int ne_row, int n_edge, int nseg,
                              const float *hr, const float *hi, float *q_r, float *q_i) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_edge) return;
    int iy, ix, iz;
    qedge_point(j, Nx, Ny, Nz, ne_row, iy, ix, iz);
    const int M2 = 2 * Nx, M3 = Nz > 1 ? 2 * Nz : 1;
    const long nq = 2L * Ny * M2 * M3;
    const long g = ((long)iy * M2 + ix) * M3 + iz;
    for (int l = 0; l < nseg; ++l) {
        q_r[(size_t)l * nq + g] = hr[(size_t)l * n_edge + j];
        q_i[(size_t)l * nq + g] = hi[(size_t)l * n_edge + j];
    }
}

using QsFn = void (*)(const float4 *, const float4 *, const float *, const half8 *,
                      const QsJob *, int, int, int, int, int, int, int, int, int, float *, float);

QsFn pick_qsep(int wm, bool d3) {
    switch (wm) {
        case 1: return d3 ? (QsFn)k_q_sep<1, true> : (QsFn)k_q_sep<1, false>;
        case 2: return d3 ? (QsFn)k_q_sep<2, true> : (QsFn)k_q_sep<2, false>;
        case 3: return d3 ? (QsFn)k_q_sep<3, true> : (QsFn)k_q_sep<3, false>;
        default: return d3 ? (QsFn)k_q_sep<4, true> : (QsFn)k_q_sep<4, false>;
    }
}

// Standard-grid Q by the separable GEMM.  W = the k_w_q weights
// [nseg][num_k]; spos = the direct kernel's sample stream (edge points).
int toeplitz_q_sep(imp_ctx *ctx, int num_k, const float *kx, const float *ky, const float *kz,
                   const float *t, const double *spos, const float2 *W, bool d3, int Nx, int Ny,
                   int Nz, int nseg, float tau, float t0, bool cumulative, float *q_r,
                   float *q_i) {
    const hipStream_t st = ctx->stream;
    const int M2 = 2 * Nx, M3 = d3 ? 2 * Nz : 1;
    const int nbk = nseg + 1;
    const int nblk = ceil_div(num_k, QS_BLK);
    // bucketed positions: at most num_k + nbk * 7, plus up to two zero steps
    const long kp_max = (long)num_k + (long)(nbk + 2) * MF_K;
    // 1. bucket the samples by segment (stable counting sort)
    void *sw = nullptr;
    const size_t n_cnt = (size_t)nbk * nblk, n_bs = 64;
    IMP_TRY(workspace(ctx, WS_QS_SORT, (n_cnt + n_bs + 64 + (size_t)kp_max) * sizeof(int), &sw));
    int *cnt = (int *)sw, *bstart = cnt + n_cnt;
    unsigned int *maxbits = (unsigned int *)(bstart + n_bs);
    int *perm = bstart + n_bs + 64;
    hipLaunchKernelGGL(k_fill_i32, dim3(ceil_div(kp_max, 256)), dim3(256), 0, st, perm, kp_max,
                       -1);
    hipLaunchKernelGGL(k_qb_count, dim3(nblk), dim3(256), 0, st, num_k, t, tau, t0, nseg, nblk,
                       cnt);
    hipLaunchKernelGGL(k_qb_scan, dim3(1), dim3(1024), 0, st, nblk, nbk, cnt, bstart);
    hipLaunchKernelGGL(k_qb_scatter, dim3(nblk), dim3(256), 0, st, num_k, t, tau, t0, nseg, nblk,
                       (const int *)cnt, perm);
    IMP_HIP(hipMemsetAsync(maxbits, 0, sizeof(unsigned int), st));
    hipLaunchKernelGGL(k_absmax, dim3(512), dim3(256), 0, st, W, (size_t)nseg * num_k, maxbits);
    IMP_CHECK_LAUNCH();
    std::vector<int> bs(nbk + 1);
    IMP_HIP(hipMemcpyAsync(bs.data(), bstart, (nbk + 1) * sizeof(int), hipMemcpyDeviceToHost, st));
    IMP_HIP(hipStreamSynchronize(st));
    // tables hold an even number of steps with at least one zero step after
    // the last bucket: a job extended to even length (below) stays inside
    const int n_steps = (bs[nbk] / MF_K + 2) & ~1;
    const int kp = n_steps * MF_K;
    // 2. jobs: segment l runs over buckets l and l + 1, split into K blocks
    //    so the launch holds ~16 workgroups per CU
    const int R = Ny * M3;
    const int wm = R >= 256 ? 4 : ceil_div(R, 64);
    const int ntx = ceil_div(M2, 16);
    const int gx = ceil_div(ntx, QS_TILES), gy = ceil_div(R, 64 * wm);
    const int ntp = gx * QS_TILES;
    long tot_steps = 0;
    for (int l = 0; l < nseg; ++l) tot_steps += (bs[l + 2] - bs[l]) / MF_K;
    const long target = 16L * ctx->num_cus;
    long kb = tot_steps * gx * gy / std::max(1L, target);
    kb = std::min(2048L, std::max((long)MF_FOLD * 2, (kb + MF_FOLD - 1) / MF_FOLD * MF_FOLD));
    // K-block partials stay within QS_PART_CAP: on large grids fewer, longer
    // blocks (the two-sum folds keep their accuracy over any length)
    const size_t part_per_job = (size_t)2 * R * M2 * sizeof(float);
    const long max_jobs = std::max<long>(2L * nseg, (long)(QS_PART_CAP / part_per_job));
    if (tot_steps / kb + nseg > max_jobs) {
        kb = (tot_steps + (max_jobs - nseg) - 1) / (max_jobs - nseg);
        kb = (kb + MF_FOLD - 1) / MF_FOLD * MF_FOLD;
    }
    std::vector<QsJob> jobs;
    std::vector<int> jb(nseg + 1, 0);
    for (int l = 0; l < nseg; ++l) {
        jb[l] = (int)jobs.size();
        const int a = bs[l] / MF_K, b = bs[l + 2] / MF_K;
        // every job an even number of steps (kb is): a last odd one takes
        // one more step, whose samples enter with their own weights a_l
        // (zero beyond the two buckets) or are zero padding
        for (int s = a; s < b; s += (int)kb) {
            int e = std::min(b, s + (int)kb);
            if ((e - s) & 1) ++e;
            jobs.push_back(QsJob{l, s, e, 0});
        }
    }
    jb[nseg] = (int)jobs.size();
    const int njobs = (int)jobs.size();
    for (const QsJob &j : jobs)
        if (j.s0 < 0 || j.s1 > n_steps || j.s1 - j.s0 < 2 || ((j.s1 - j.s0) & 1))
            return fail(IMP_EINVAL, "toeplitz_q: K block [%d, %d) outside %d steps", j.s0, j.s1,
                        n_steps);
    if ((long)njobs * gx * gy > 0x7fffffffL)
        return fail(IMP_EUNSUPPORTED, "toeplitz_q: %d K blocks", njobs);
    // 3. tables: wq [nseg][kp], ey [step][Ny][8], ez [step][M3][8], bx
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t b_wq = al((size_t)nseg * kp * sizeof(float));
    const size_t b_ey = al((size_t)kp * Ny * sizeof(float2));
    const size_t b_ez = d3 ? al((size_t)kp * M3 * sizeof(float2)) : 0;
    const size_t b_bx = al((size_t)n_steps * ntp * 2 * 64 * sizeof(half8));
    if (b_wq + b_ey + b_ez + b_bx > QS_TAB_CAP) return QS_FALLBACK;
    void *tw = nullptr;
    IMP_TRY(workspace(ctx, WS_QS_TAB, b_wq + b_ey + b_ez + b_bx + 256, &tw));
    float *wq = (float *)tw;
    float2 *ey = (float2 *)((char *)tw + b_wq);
    float2 *ez = (float2 *)((char *)tw + b_wq + b_ey);
    half8 *bx = (half8 *)((char *)tw + b_wq + b_ey + b_ez);
    if (kp > 0) {
        hipLaunchKernelGGL(k_qt_w, dim3(ceil_div((long)kp * nseg, 256)), dim3(256), 0, st, kp,
                           nseg, num_k, (const int *)perm, W, (const unsigned int *)maxbits, wq);
        hipLaunchKernelGGL(k_qt_line, dim3(ceil_div((long)kp * Ny, 256)), dim3(256), 0, st, kp,
                           Ny, Ny, (const int *)perm, ky, ey);
        if (d3)
            hipLaunchKernelGGL(k_qt_line, dim3(ceil_div((long)kp * M3, 256)), dim3(256), 0, st,
                               kp, M3, Nz, (const int *)perm, kz, ez);
        hipLaunchKernelGGL(k_qt_bx, dim3(n_steps, ceil_div(ntp, 16)), dim3(256), 0, st, ntp, M2,
                           Nx, (const int *)perm, kx, bx);
        IMP_CHECK_LAUNCH();
    }
    // 4. GEMM over K blocks, FP64 reduce with the mirror fill
    const size_t plane = (size_t)R * M2;
    const size_t b_part = al((size_t)std::max(1, njobs) * 2 * plane * sizeof(float));
    void *pw = nullptr;
    IMP_TRY(workspace(ctx, WS_QS_PART, b_part + al(njobs * sizeof(QsJob)) + al(jb.size() * 4),
                      &pw));
    float *part = (float *)pw;
    QsJob *djobs = (QsJob *)((char *)pw + b_part);
    int *djb = (int *)((char *)pw + b_part + al(njobs * sizeof(QsJob)));
    if (njobs > 0)
        IMP_HIP(hipMemcpyAsync(djobs, jobs.data(), njobs * sizeof(QsJob), hipMemcpyHostToDevice,
                               st));
    IMP_HIP(hipMemcpyAsync(djb, jb.data(), jb.size() * sizeof(int), hipMemcpyHostToDevice, st));
    if (njobs > 0) {
        const int rec = prof_begin(ctx, "k_q_sep");
        const long nwg = (long)ceil_div(njobs, 8) * 8 * gx * gy;
        hipLaunchKernelGGL(pick_qsep(wm, d3), dim3((unsigned)nwg), dim3(64 * wm), 0, st,
                           (const float4 *)ey, (const float4 *)ez, (const float *)wq,
                           (const half8 *)bx, (const QsJob *)djobs, njobs, gx, gy, R, Ny, M3,
                           ntp, kp, M2, part, 1.0f);
        prof_end(ctx, rec);
        IMP_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_qs_reduce, dim3(ceil_div((long)plane, 256)), dim3(256), 0, st,
                       (const float *)part, (const int *)djb, nseg, R, M2, Nx, Ny, Nz, M3,
                       cumulative ? 1 : 0, (const unsigned int *)maxbits, q_r, q_i);
    IMP_CHECK_LAUNCH();
    // 5. the row y = 0 and the mirror-less points of rows y > 0, directly
    const int ne_row = M3 > 1 ? M2 + M3 - 1 : 1;
    const int n_edge = M2 * M3 + (Ny - 1) * ne_row;
    void *lanes = nullptr, *half = nullptr;
    IMP_TRY(workspace(ctx, WS_LANES, (size_t)n_edge * sizeof(float4), &lanes));
    hipLaunchKernelGGL(k_lanes_qedge, dim3(ceil_div(n_edge, 256)), dim3(256), 0, st, Nx, Ny,
                       d3 ? Nz : 1, ne_row, n_edge, (float4 *)lanes);
    IMP_CHECK_LAUNCH();
    IMP_TRY(workspace(ctx, WS_QHALF, 2 * (size_t)nseg * n_edge * sizeof(float), &half));
    float *h_r = (float *)half, *h_i = h_r + (size_t)nseg * n_edge;
    Job J{(const float4 *)lanes, n_edge, spos, num_k, W, nseg, d3, +1.f,
          cumulative ? Out::PerRhsCumulative : Out::PerRhs, nullptr, nullptr, h_r, h_i};
    IMP_TRY(run_job(ctx, J));
    hipLaunchKernelGGL(k_q_edge_fill, dim3(ceil_div(n_edge, 256)), dim3(256), 0, st, Nx, Ny,
                       d3 ? Nz : 1, ne_row, n_edge, nseg, (const float *)h_r, (const float *)h_i,
                       q_r, q_i);
    IMP_CHECK_LAUNCH();
    // the host job table must outlive its copy; large tables are not kept
    IMP_HIP(hipStreamSynchronize(st));
    IMP_TRY(workspace_release(ctx, WS_QS_TAB, QS_KEEP));
    IMP_TRY(workspace_release(ctx, WS_QS_PART, QS_KEEP));
    return IMP_OK;
}

}  // namespace

namespace {

// Q_l of a stack on the standard 3D grid: the Q weights a_l(t) |phi|^2 and
// kx, ky repeat in every partition, so
//   Q_l(x, y, z) = Q2_l(x, y) Z(z),  Z(z) = sum_p e^{+i 2pi kz_p z},
// with Q2_l the 2D kernel of one partition.  Z is evaluated in FP64 at the
// grid's z = (iz - Nz) / Nz (the FP32 value the direct kernel uses).
__global__ void k_qz_factor(int P, int K, const float *kz, int Nz, double2 *Z) {
    const int iz = blockIdx.x * blockDim.x + threadIdx.x;
    if (iz >= 2 * Nz) return;
    const double z = (double)((float)(iz - Nz) / (float)Nz);
    double zr = 0.0, zi = 0.0;
    for (int p = 0; p < P; ++p) {
        double s, c;
        sincospi(2.0 * ((double)kz[(size_t)p * K] * z), &s, &c);
        zr += c;
        zi += s;
    }
    Z[iz] = make_double2(zr, zi);
}

// q[l][(iy M2 + ix) M3 + iz] = q2[l][iy M2 + ix] * Z[iz]
__global__ void k_q_outer(int nseg, long n2, int M3, const float *q2r, const float *q2i,
                          const double2 *Z, float *q_r, float *q_i) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nq = n2 * M3;
    if (g >= nq) return;
    const int iz = (int)(g % M3);
    const long p2 = g / M3;
    const double2 z = Z[iz];
    for (int l = 0; l < nseg; ++l) {
        const double ar = q2r[(size_t)l * n2 + p2], ai = q2i[(size_t)l * n2 + p2];
        q_r[(size_t)l * nq + g] = (float)(ar * z.x - ai * z.y);
        q_i[(size_t)l * nq + g] = (float)(ar * z.y + ai * z.x);
    }
}

// Standard grid, y outer, x, z inner (y = iy - Ny, x = ix - Nx, z = iz - Nz):
// Q(-r) = conj(Q(r)) exactly.  The mirror-filled paths compute the row
// y = 0 point by point, so its pairs (x, 0, z), (-x, 0, -z) agree only to
// rounding; here every point after the origin in (y, x, z) order whose
// mirror is on the grid takes the conjugate of that mirror, and Q(0) is
// real.  (Points with y = -Ny, x = -Nx or z = -Nz have no mirror.)
__global__ void k_q_conj_sym(int nseg, int Nx, int Ny, int Nz, long nq, float *q_r,
                             float *q_i) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nq) return;
    const int M2 = 2 * Nx, M3 = Nz > 1 ? 2 * Nz : 1;
    const int iz = (int)(g % M3);
    const int ix = (int)((g / M3) % M2);
    const int iy = (int)(g / ((long)M3 * M2));
    if (iy == 0 || ix == 0 || (M3 > 1 && iz == 0)) return;
    const int y = iy - Ny, x = ix - Nx, z = M3 > 1 ? iz - Nz : 0;
    if (y == 0 && x == 0 && z == 0) {
        for (int l = 0; l < nseg; ++l) q_i[(size_t)l * nq + g] = 0.f;
        return;
    }
    const bool upper = y > 0 || (y == 0 && (x > 0 || (x == 0 && z > 0)));
    if (!upper) return;
    const long m = ((long)(2 * Ny - iy) * M2 + (2 * Nx - ix)) * M3 + (M3 > 1 ? 2 * Nz - iz : 0);
    for (int l = 0; l < nseg; ++l) {
        q_r[(size_t)l * nq + g] = q_r[(size_t)l * nq + m];
        q_i[(size_t)l * nq + g] = -q_i[(size_t)l * nq + m];
    }
}

int q_conj_sym(imp_ctx *ctx, int nseg, int Nx, int Ny, int Nz, float *q_r, float *q_i) {
    const long nq = 4L * Nx * Ny * (Nz > 1 ? 2L * Nz : 1L);
    hipLaunchKernelGGL(k_q_conj_sym, dim3(ceil_div(nq, 256)), dim3(256), 0, ctx->stream, nseg,
                       Nx, Ny, Nz, nq, q_r, q_i);
    IMP_CHECK_LAUNCH();
    return IMP_OK;
}

}  // namespace

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
    if (std_grid && Nz > 1 && !kz)
        return fail(IMP_EINVAL, "toeplitz_q: a 3D grid (Nz = %d) needs kz", Nz);
    const bool d3 = kz && (std_grid ? Nz > 1 : gz != nullptr);
    const int nseg = ntime_segs + 1;
    if (nseg > 32) return fail(IMP_EUNSUPPORTED, "toeplitz_q: at most 31 time segments");
    // a stack on the standard grid: the 2D kernel of one partition times Z(z)
    if (std_grid && d3 && !(flags & IMP_Q_DIRECT)) {
        int K = 0;
        IMP_TRY(find_stack(ctx, num_k, kx, ky, kz, t, phi_r, phi_i, &K));
        if (K > 0) {
            const long n2 = 4L * Nx * Ny;
            const size_t planes = (size_t)nseg * n2;
            void *buf = nullptr;
            IMP_TRY(workspace(ctx, WS_QSTACK,
                              2 * planes * sizeof(float) + 2 * Nz * sizeof(double2), &buf));
            float *q2r = (float *)buf, *q2i = q2r + planes;
            double2 *Z = (double2 *)(q2i + planes);
            IMP_TRY(toeplitz_q(ctx, K, kx, ky, nullptr, t, phi_r, phi_i, Nx, Ny, 1, (int)n2,
                               nullptr, nullptr, nullptr, ntime_segs, flags, t_host_first_last,
                               q2r, q2i));
            hipLaunchKernelGGL(k_qz_factor, dim3(ceil_div(2 * Nz, 64)), dim3(64), 0, ctx->stream,
                               num_k / K, K, kz, Nz, Z);
            IMP_CHECK_LAUNCH();
            const int rec = prof_begin(ctx, "k_q_outer");
            hipLaunchKernelGGL(k_q_outer, dim3(ceil_div(n2 * 2 * Nz, 256)), dim3(256), 0,
                               ctx->stream, nseg, n2, 2 * Nz, (const float *)q2r,
                               (const float *)q2i, (const double2 *)Z, q_r, q_i);
            prof_end(ctx, rec);
            IMP_CHECK_LAUNCH();
            return q_conj_sym(ctx, nseg, Nx, Ny, Nz, q_r, q_i);
        }
    }
    // tau and t0 as the reference computes them, in FP32 on the host
    // (computeQ.cmem.cu:870-871)
    const float t0 = t_host_first_last[0];
    const float tlast = t_host_first_last[1];
    const float tau = ntime_segs > 0 ? (tlast - t0) / (float)ntime_segs : (tlast - t0);
    const double *spos = nullptr;
    IMP_TRY(build_spos(ctx, num_k, kx, ky, d3 ? kz : nullptr, nullptr, 0.0, &spos));
    void *lanes = nullptr, *W = nullptr;
    IMP_TRY(workspace(ctx, WS_W, (size_t)num_k * nseg * sizeof(float2), &W));
    hipLaunchKernelGGL(k_w_q, dim3(ceil_div((long)num_k * nseg, 256)), dim3(256), 0,
                       ctx->stream, num_k, nseg, t, phi_r, phi_i, tau, t0, (float2 *)W);
    IMP_CHECK_LAUNCH();
    const Out mode = (flags & IMP_Q_CUMULATIVE) ? Out::PerRhsCumulative : Out::PerRhs;
    // Standard grid (main_mri.cu:461-500; y outer, x, z inner, r = (x/Nx,
    // y/Ny, z/Nz)): the separable GEMM unless the direct kernel is asked for
    // (or tau is degenerate); the direct kernel evaluates the half grid and
    // fills the rest by Hermitian symmetry.
    if (std_grid && !(flags & IMP_Q_DIRECT) && tau > 0.f && std::isfinite(tau) &&
        std::isfinite(t0)) {
        const int rc = toeplitz_q_sep(ctx, num_k, kx, ky, d3 ? kz : nullptr, t, spos,
                                      (const float2 *)W, d3, Nx, Ny, Nz, nseg, tau, t0,
                                      mode == Out::PerRhsCumulative, q_r, q_i);
        if (rc == IMP_OK) return q_conj_sym(ctx, nseg, Nx, Ny, Nz, q_r, q_i);
        if (rc != QS_FALLBACK) return rc;
    }
    const int M2 = 2 * Nx, M3 = Nz > 1 ? 2 * Nz : 1;
    const int n1 = std_grid ? (Ny + 1) * M2 * M3 : 0;
    const int ne_row = M3 > 1 ? M2 + M3 - 1 : 1;
    const int n_comp = std_grid ? n1 + (Ny - 1) * ne_row : num_q;
    IMP_TRY(workspace(ctx, WS_LANES, (size_t)n_comp * sizeof(float4), &lanes));
    if (std_grid) {
        hipLaunchKernelGGL(k_lanes_qhalf, dim3(ceil_div(n_comp, 256)), dim3(256), 0, ctx->stream,
                           Nx, Ny, Nz, n1, ne_row, n_comp, (float4 *)lanes);
    } else {
        hipLaunchKernelGGL(k_lanes_from_arrays, dim3(ceil_div(num_q, 256)), dim3(256), 0,
                           ctx->stream, num_q, gx, gy, d3 ? gz : nullptr, nullptr, 0.0,
                           (float4 *)lanes);
    }
    IMP_CHECK_LAUNCH();
    if (!std_grid) {
        Job J{(const float4 *)lanes, num_q, spos, num_k, (const float2 *)W, nseg, d3, +1.f,
              mode, nullptr, nullptr, q_r, q_i};
        return run_job(ctx, J);
    }
    void *half = nullptr;
    IMP_TRY(workspace(ctx, WS_QHALF, 2 * (size_t)nseg * n_comp * sizeof(float), &half));
    float *h_r = (float *)half, *h_i = h_r + (size_t)nseg * n_comp;
    Job J{(const float4 *)lanes, n_comp, spos, num_k, (const float2 *)W, nseg, d3, +1.f,
          mode, nullptr, nullptr, h_r, h_i};
    IMP_TRY(run_job(ctx, J));
    hipLaunchKernelGGL(k_q_fill, dim3(ceil_div(num_q, 256)), dim3(256), 0, ctx->stream, Nx, Ny,
                       Nz, n1, ne_row, n_comp, nseg, (const float *)h_r, (const float *)h_i, q_r,
                       q_i);
    IMP_CHECK_LAUNCH();
    return q_conj_sym(ctx, nseg, Nx, Ny, Nz, q_r, q_i);
}

}  // namespace imp

extern "C" int imp_nudft_adjoint_batch(imp_ctx *ctx, int njobs, const imp_nudft_job *jobs,
                                       int chunk_scale) {
    return imp::nudft_adjoint_batch(ctx, njobs, jobs, chunk_scale);
}

extern "C" int imp_nudft_adjoint_noalloc(imp_ctx *ctx, int num_k, int num_x, int ncoils,
                                         const float *kx, const float *ky, const float *kz,
                                         const float *t, const float *x, const float *y,
                                         const float *z, const float *fm, const float *phi_r,
                                         const float *phi_i, const float *d_r, const float *d_i,
                                         const float *s_r, const float *s_i, float *out_r,
                                         float *out_i) {
    return imp::nudft_adjoint_noalloc(ctx, num_k, num_x, ncoils, kx, ky, kz, t, x, y, z, fm,
                                      phi_r, phi_i, d_r, d_i, s_r, s_i, out_r, out_i);
}
