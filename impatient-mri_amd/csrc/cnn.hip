// Learned warm start: a 6-layer 3x3 residual CNN in BF16 with FP32
// accumulation, with its own inference and training kernels.
//
// Every activation is an H x W x 32 image (NHWC, channels contiguous, bf16);
// the 2-channel input (Re, Im of the normalised linear estimate) and output
// use channels 0-1 of the same layout, so all six layers run one 32 -> 32
// convolution kernel.  Layers 0-4 are conv + bias + ReLU, layer 5 is conv +
// bias (FP32 out); the network output is added to its input (residual).
//
// Convolution as implicit GEMM on v_mfma_f32_32x32x16_bf16: a wave computes
// 32 pixels of one row x 32 output channels; K runs over (tap, input
// channel) in 18 chunks of 16.  A operand = 8 consecutive input channels of
// one pixel (one 16-byte load), B operand = 8 consecutive input channels of
// one output channel of the tap ([tap][co][ci] weight layout).  The backward
// data pass is the same kernel with the taps flipped and co/ci swapped, the
// ReLU mask fused into its store; the weight gradient is a second MFMA
// kernel that stages 64-pixel row tiles (three rows, three x shifts) in LDS
// so both operands are 16-byte aligned reads.
//
// No counterpart in the reference (its CG always starts from zero,
// recon.cu:1871, 2011-2133).
#include "common.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace imp {

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int CH = 32;                 // channels of every activation
constexpr int NLAYER = 6;
constexpr int WSZ = 9 * CH * CH;       // weights per layer, [tap][co][ci]
constexpr int PSZ = WSZ + CH;          // weights + bias per layer
constexpr int NPARAM = NLAYER * PSZ;
constexpr int CIO = 2;                 // real channels of network input / output
constexpr int WG_TILE = 64;            // weight-gradient pixels per tile
constexpr int WG_BLOCKS = 512;         // weight-gradient workgroups (partials)
constexpr int RED_T = 256;

// ------------------------------------------------------------ convolution
enum { CONV_RELU = 0, CONV_LINEAR = 1, CONV_MASK = 2 };

// out[p][co] = bias[co] + sum_{tap, ci} in[p + off(tap)][ci] * w[tap][co][ci]
// (zero padding), tap = (dy + 1) * 3 + (dx + 1).  W % 32 == 0.
//   CONV_RELU:   out bf16 = max(v, 0)
//   CONV_LINEAR: out fp32 = v
//   CONV_MASK:   out bf16 = v if mask[p][co] > 0 else 0   (bias unused)
template <int MODE>
__global__ __launch_bounds__(256) void k_conv3x3(const bf16 *__restrict__ in,
                                                 const bf16 *__restrict__ w,
                                                 const float *__restrict__ bias,
                                                 const bf16 *__restrict__ mask,
                                                 void *__restrict__ out, int H, int W) {
    const int l = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tiles_per_row = W / 32;
    const int tile = blockIdx.x * 4 + wave;
    if (tile >= H * tiles_per_row) return;  // no barriers below
    const int y = tile / tiles_per_row, x0 = (tile % tiles_per_row) * 32;
    const int m = l & 31, h = l >> 5;
    floatx16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;
    const bf16x8 zero = {};
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
        const int yy = y + tap / 3 - 1, xx = x0 + m + tap % 3 - 1;
        const bool ok = yy >= 0 && yy < H && xx >= 0 && xx < W;
        const bf16x8 *src =
            (const bf16x8 *)(in + ((size_t)(ok ? yy : 0) * W + (ok ? xx : 0)) * CH);
        const bf16x8 *wp = (const bf16x8 *)(w + ((size_t)tap * CH + m) * CH);
#pragma unroll
        for (int kc = 0; kc < 2; ++kc) {
            const bf16x8 a = ok ? src[kc * 2 + h] : zero;  // ci = 16 kc + 8 h + j
            const bf16x8 b = wp[kc * 2 + h];               // co = m, same ci
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
    }
    // D layout: lane l, register i -> column l % 32 (co), row 8(i/4) + 4h + i%4 (pixel)
    const int co = l & 31;
    const float bv = (MODE != CONV_MASK && bias) ? bias[co] : 0.f;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int px = x0 + 8 * (i >> 2) + 4 * h + (i & 3);
        const size_t o = ((size_t)y * W + px) * CH + co;
        const float v = acc[i] + bv;
        if constexpr (MODE == CONV_RELU) {
            ((bf16 *)out)[o] = (bf16)fmaxf(v, 0.f);
        } else if constexpr (MODE == CONV_LINEAR) {
            ((float *)out)[o] = v;
        } else {
            ((bf16 *)out)[o] = (float)mask[o] > 0.f ? (bf16)v : (bf16)0.f;
        }
    }
}

// ------------------------------------------------------- weight gradient
// part[block][tap][co][ci] = sum over the block's pixels p of
//   g[p][co] * a[p + off(tap)][ci];  part[block][WSZ + co] = sum g[p][co].
// 64-pixel row tiles (W % 64 == 0), grid-stride over tiles.  LDS holds
// aS[tap][ci][p] = a[y + dy][x0 + p + dx][ci] (zero outside) and
// gS[co][p] = g[y][x0 + p][co]; wave w accumulates taps w, w + 4, w + 8 with
// D[co][ci] += A[co][pixels] B[pixels][ci].
__global__ __launch_bounds__(256) void k_conv_wgrad(const bf16 *__restrict__ a,
                                                    const bf16 *__restrict__ g, int H, int W,
                                                    float *__restrict__ part) {
    __shared__ __attribute__((aligned(16))) bf16 aS[9][CH][WG_TILE];
    __shared__ __attribute__((aligned(16))) bf16 gS[CH][WG_TILE];
    __shared__ float bS[8][CH];
    const int t = threadIdx.x;
    const int l = t & 63, wave = t >> 6, h = l >> 5;
    floatx16 acc[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[k][i] = 0.f;
    float bacc = 0.f;
    const int tiles_per_row = W / WG_TILE, ntiles = H * tiles_per_row;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int y = tile / tiles_per_row, x0 = (tile % tiles_per_row) * WG_TILE;
        __syncthreads();  // the previous tile's LDS reads are done
        // g tile: 64 pixels x 4 chunks of 8 channels
        {
            const int p = t >> 2, c0 = (t & 3) * 8;
            const bf16x8 v = *(const bf16x8 *)(g + ((size_t)y * W + x0 + p) * CH + c0);
#pragma unroll
            for (int j = 0; j < 8; ++j) gS[c0 + j][p] = v[j];
        }
        // a rows y-1..y+1, source pixels x0-1 .. x0+64 (66), 4 chunks each
        for (int e = t; e < 3 * (WG_TILE + 2) * 4; e += 256) {
            const int dyi = e / ((WG_TILE + 2) * 4);
            const int rem = e % ((WG_TILE + 2) * 4);
            const int q = rem / 4 - 1;  // source pixel offset from x0, -1..64
            const int c0 = (rem % 4) * 8;
            const int yy = y + dyi - 1, xx = x0 + q;
            bf16x8 v = {};
            if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                v = *(const bf16x8 *)(a + ((size_t)yy * W + xx) * CH + c0);
#pragma unroll
            for (int dxi = 0; dxi < 3; ++dxi) {
                const int p = q - (dxi - 1);  // a[y+dy][x0 + p + dx] with dx = dxi - 1
                if (p >= 0 && p < WG_TILE) {
#pragma unroll
                    for (int j = 0; j < 8; ++j) aS[dyi * 3 + dxi][c0 + j][p] = v[j];
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int tap = wave + 4 * k;
            if (tap < 9) {
#pragma unroll
                for (int kc = 0; kc < WG_TILE / 16; ++kc) {
                    const bf16x8 A = *(const bf16x8 *)&gS[l & 31][kc * 16 + 8 * h];
                    const bf16x8 B = *(const bf16x8 *)&aS[tap][l & 31][kc * 16 + 8 * h];
                    acc[k] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, B, acc[k], 0, 0, 0);
                }
            }
        }
        {  // bias: co = t % 32, pixels t/32 + 8 j
            const int co = t & 31;
#pragma unroll
            for (int j = 0; j < WG_TILE / 8; ++j) bacc += (float)gS[co][(t >> 5) + 8 * j];
        }
    }
    float *dst = part + (size_t)blockIdx.x * PSZ;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int tap = wave + 4 * k;
        if (tap < 9) {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int co = 8 * (i >> 2) + 4 * h + (i & 3), ci = l & 31;
                dst[((size_t)tap * CH + co) * CH + ci] = acc[k][i];
            }
        }
    }
    bS[t >> 5][t & 31] = bacc;
    __syncthreads();
    if (t < CH) {
        float s = 0.f;
#pragma unroll
        for (int r = 0; r < 8; ++r) s += bS[r][t];
        dst[WSZ + t] = s;
    }
}

// grad[i] = sum_b part[b][i]
__global__ void k_sum_parts(const float *__restrict__ part, int nparts, float *__restrict__ grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= PSZ) return;
    float s = 0.f;
    for (int b = 0; b < nparts; ++b) s += part[(size_t)b * PSZ + i];
    grad[i] = s;
}

// ------------------------------------------------------------ parameters
// trainable mask: layer 0 only reads input channels < CIO, layer 5 only
// writes output channels < CIO; the other weights stay zero.
__device__ __forceinline__ bool trainable(int idx) {
    const int layer = idx / PSZ, r = idx % PSZ;
    if (r >= WSZ) return layer != NLAYER - 1 || (r - WSZ) < CIO;
    const int co = (r / CH) % CH, ci = r % CH;
    if (layer == 0 && ci >= CIO) return false;
    if (layer == NLAYER - 1 && co >= CIO) return false;
    return true;
}

// master FP32 -> bf16 forward weights [tap][co][ci] and backward-data
// weights wb[tap'][ci][co] = w[8 - tap'][co][ci]
__global__ void k_weights_bf16(const float *__restrict__ params, bf16 *__restrict__ wf,
                               bf16 *__restrict__ wb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NLAYER * WSZ) return;
    const int layer = i / WSZ, r = i % WSZ;
    const int tap = r / (CH * CH), co = (r / CH) % CH, ci = r % CH;
    const bf16 v = (bf16)params[(size_t)layer * PSZ + r];
    wf[i] = v;
    wb[(size_t)layer * WSZ + ((size_t)(8 - tap) * CH + ci) * CH + co] = v;
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ grad,
                       float *__restrict__ m, float *__restrict__ v, float lr, float b1, float b2,
                       float eps, float bc1, float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    if (!trainable(i)) {
        p[i] = 0.f;
        return;
    }
    const float g = grad[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// ------------------------------------------------------ input / output
__global__ __launch_bounds__(RED_T) void k_sumsq(const float *__restrict__ re,
                                                 const float *__restrict__ im, long n,
                                                 double *__restrict__ partials) {
    double s = 0.0;
    for (long i = (long)blockIdx.x * RED_T + threadIdx.x; i < n; i += (long)gridDim.x * RED_T)
        s += (double)re[i] * re[i] + (double)im[i] * im[i];
    __shared__ double sm[RED_T];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int k = RED_T / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) sm[threadIdx.x] += sm[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = sm[0];
}

__global__ void k_finish_sum(const double *partials, int np, double *out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < np; ++i) s += partials[i];
        *out = s;
    }
}

// a0[p] = (z[p] / s, 0, ..., 0) in bf16, s = sqrt(sumsq / n) (sumsq on device)
__global__ void k_features(const float *__restrict__ zr, const float *__restrict__ zi, long n,
                           const double *__restrict__ sumsq, bf16 *__restrict__ a0) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double ss = *sumsq;
    const float inv = ss > 0.0 ? (float)(1.0 / sqrt(ss / (double)n)) : 0.f;
    bf16 *o = a0 + p * CH;
    o[0] = (bf16)(zr[p] * inv);
    o[1] = (bf16)(zi[p] * inv);
#pragma unroll
    for (int c = 2; c < CH; ++c) o[c] = (bf16)0.f;
}

// x = z + s * net[p][0..1]  (= s * (u + net) with u = z / s)
__global__ void k_output(const float *__restrict__ zr, const float *__restrict__ zi,
                         const float *__restrict__ net, long n, const double *__restrict__ sumsq,
                         float *__restrict__ xr, float *__restrict__ xi) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float s = (float)sqrt(*sumsq / (double)n);
    xr[p] = zr[p] + s * net[p * CH + 0];
    xi[p] = zi[p] + s * net[p * CH + 1];
}

// loss 0.5 |u + net - t/s|^2 / |t/s|^2 per pixel (partials) and its
// gradient with respect to net, in bf16 (channels >= CIO zero)
__global__ __launch_bounds__(RED_T) void k_loss_grad(
    const float *__restrict__ zr, const float *__restrict__ zi, const float *__restrict__ tr,
    const float *__restrict__ ti, const float *__restrict__ net, long n,
    const double *__restrict__ sums /* [0] sumsq z, [1] sumsq t */, bf16 *__restrict__ g,
    double *__restrict__ partials) {
    const double ssz = sums[0], sst = sums[1];
    const double s = sqrt(ssz / (double)n);
    const float inv = s > 0.0 ? (float)(1.0 / s) : 0.f;
    const double tnorm = s > 0.0 ? sst / (s * s) : 1.0;  // |t/s|^2
    const float scale = tnorm > 0.0 ? (float)(1.0 / tnorm) : 0.f;
    double acc = 0.0;
    for (long p = (long)blockIdx.x * RED_T + threadIdx.x; p < n; p += (long)gridDim.x * RED_T) {
        const float er = zr[p] * inv + net[p * CH + 0] - tr[p] * inv;
        const float ei = zi[p] * inv + net[p * CH + 1] - ti[p] * inv;
        acc += 0.5 * ((double)er * er + (double)ei * ei);
        bf16 *o = g + p * CH;
        o[0] = (bf16)(er * scale);
        o[1] = (bf16)(ei * scale);
#pragma unroll
        for (int c = 2; c < CH; ++c) o[c] = (bf16)0.f;
    }
    __shared__ double sm[RED_T];
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int k = RED_T / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) sm[threadIdx.x] += sm[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = sm[0] * scale;
}

// deterministic He-normal initialisation (splitmix64 + Box-Muller), host side
uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
double unif(uint64_t &s) { return ((splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

}  // namespace
}  // namespace imp

using namespace imp;

struct imp_cnn {
    imp_ctx *ctx = nullptr;
    int H = 0, W = 0;
    long n = 0;
    float *params = nullptr, *grad = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    bf16 *wf = nullptr, *wb = nullptr;
    bf16 *act[NLAYER] = {};   // act[0] = features, act[l] = input of layer l
    float *net = nullptr;     // layer-5 output [n][CH]
    bf16 *g0 = nullptr, *g1 = nullptr;
    float *part = nullptr;
    double *dsc = nullptr;    // [0..RED_T) partials, then sums
    int step = 0;
    bool weights_dirty = true;
    std::vector<void *> owned;
};

namespace {

int cnn_alloc(imp_cnn *net, size_t bytes, void **p) {
    if (hipMalloc(p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return fail(IMP_ENOMEM, "imp_cnn: cannot allocate %zu bytes", bytes);
    }
    net->owned.push_back(*p);
    return IMP_OK;
}

int sync_weights(imp_cnn *net) {
    if (!net->weights_dirty) return IMP_OK;
    hipLaunchKernelGGL(k_weights_bf16, dim3(ceil_div(NLAYER * WSZ, 256)), dim3(256), 0,
                       net->ctx->stream, net->params, net->wf, net->wb);
    IMP_CHECK_LAUNCH();
    net->weights_dirty = false;
    return IMP_OK;
}

// sum of squares of (re, im) into dsc[RED_T + slot]
int sumsq(imp_cnn *net, const float *re, const float *im, int slot) {
    hipStream_t st = net->ctx->stream;
    hipLaunchKernelGGL(k_sumsq, dim3(RED_T), dim3(RED_T), 0, st, re, im, net->n, net->dsc);
    hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(1), 0, st, net->dsc, RED_T,
                       net->dsc + RED_T + slot);
    IMP_CHECK_LAUNCH();
    return IMP_OK;
}

int conv_grid(imp_cnn *net) { return ceil_div((long)net->H * (net->W / 32), 4); }

// features -> act[0]; layers -> net->net
int forward(imp_cnn *net, const float *zr, const float *zi) {
    hipStream_t st = net->ctx->stream;
    IMP_TRY(sync_weights(net));
    IMP_TRY(sumsq(net, zr, zi, 0));
    const int rec = prof_begin(net->ctx, "cnn_forward");
    hipLaunchKernelGGL(k_features, dim3(ceil_div(net->n, 256)), dim3(256), 0, st, zr, zi, net->n,
                       net->dsc + RED_T, net->act[0]);
    for (int layer = 0; layer < NLAYER; ++layer) {
        const bf16 *w = net->wf + (size_t)layer * WSZ;
        const float *b = net->params + (size_t)layer * PSZ + WSZ;
        if (layer < NLAYER - 1)
            hipLaunchKernelGGL(k_conv3x3<CONV_RELU>, dim3(conv_grid(net)), dim3(256), 0, st,
                               net->act[layer], w, b, (const bf16 *)nullptr,
                               (void *)net->act[layer + 1], net->H, net->W);
        else
            hipLaunchKernelGGL(k_conv3x3<CONV_LINEAR>, dim3(conv_grid(net)), dim3(256), 0, st,
                               net->act[layer], w, b, (const bf16 *)nullptr, (void *)net->net,
                               net->H, net->W);
    }
    prof_end(net->ctx, rec);
    IMP_CHECK_LAUNCH();
    return IMP_OK;
}

// loss and parameter gradient (net->grad) for one (z, target) pair
int backward(imp_cnn *net, const float *zr, const float *zi, const float *tr, const float *ti,
             double *loss) {
    hipStream_t st = net->ctx->stream;
    IMP_TRY(forward(net, zr, zi));
    IMP_TRY(sumsq(net, tr, ti, 1));
    hipLaunchKernelGGL(k_loss_grad, dim3(RED_T), dim3(RED_T), 0, st, zr, zi, tr, ti, net->net,
                       net->n, net->dsc + RED_T, net->g0, net->dsc);
    hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(1), 0, st, net->dsc, RED_T,
                       net->dsc + RED_T + 2);
    IMP_CHECK_LAUNCH();
    bf16 *g = net->g0, *gn = net->g1;
    for (int layer = NLAYER - 1; layer >= 0; --layer) {
        // weight + bias gradient of this layer from its input and output gradient
        hipLaunchKernelGGL(k_conv_wgrad, dim3(WG_BLOCKS), dim3(256), 0, st, net->act[layer], g,
                           net->H, net->W, net->part);
        hipLaunchKernelGGL(k_sum_parts, dim3(ceil_div(PSZ, 256)), dim3(256), 0, st, net->part,
                           WG_BLOCKS, net->grad + (size_t)layer * PSZ);
        IMP_CHECK_LAUNCH();
        if (layer == 0) break;
        // gradient into the previous layer's pre-activation (ReLU mask = act > 0)
        hipLaunchKernelGGL(k_conv3x3<CONV_MASK>, dim3(conv_grid(net)), dim3(256), 0, st, g,
                           net->wb + (size_t)layer * WSZ, (const float *)nullptr,
                           net->act[layer], (void *)gn, net->H, net->W);
        IMP_CHECK_LAUNCH();
        std::swap(g, gn);
    }
    double h = 0.0;
    IMP_HIP(hipMemcpyAsync(&h, net->dsc + RED_T + 2, sizeof(double), hipMemcpyDeviceToHost, st));
    IMP_HIP(hipStreamSynchronize(st));
    if (loss) *loss = h;
    return IMP_OK;
}

bool shape_ok(int H, int W) { return H >= 1 && W >= WG_TILE && W % WG_TILE == 0; }

}  // namespace

extern "C" int imp_cnn_create(imp_ctx *ctx, int H, int W, unsigned long long seed,
                              imp_cnn **out) {
    if (!ctx || !out) return fail(IMP_EINVAL, "imp_cnn_create: NULL argument");
    *out = nullptr;
    if (!shape_ok(H, W))
        return fail(IMP_EUNSUPPORTED, "imp_cnn_create: image %d x %d (W must be a multiple of 64)",
                    H, W);
    imp_cnn *net = new (std::nothrow) imp_cnn();
    if (!net) return fail(IMP_ENOMEM, "imp_cnn_create: host allocation");
    net->ctx = ctx;
    net->H = H;
    net->W = W;
    net->n = (long)H * W;
    int st = IMP_OK;
    void *p = nullptr;
    auto A = [&](size_t bytes) -> void * {
        if (st != IMP_OK) return nullptr;
        st = cnn_alloc(net, bytes, &p);
        return st == IMP_OK ? p : nullptr;
    };
    net->params = (float *)A(NPARAM * sizeof(float));
    net->grad = (float *)A(NPARAM * sizeof(float));
    net->adam_m = (float *)A(NPARAM * sizeof(float));
    net->adam_v = (float *)A(NPARAM * sizeof(float));
    net->wf = (bf16 *)A(NLAYER * WSZ * sizeof(bf16));
    net->wb = (bf16 *)A(NLAYER * WSZ * sizeof(bf16));
    for (int l = 0; l < NLAYER; ++l) net->act[l] = (bf16 *)A(net->n * CH * sizeof(bf16));
    net->net = (float *)A(net->n * CH * sizeof(float));
    net->g0 = (bf16 *)A(net->n * CH * sizeof(bf16));
    net->g1 = (bf16 *)A(net->n * CH * sizeof(bf16));
    net->part = (float *)A((size_t)WG_BLOCKS * PSZ * sizeof(float));
    net->dsc = (double *)A((RED_T + 8) * sizeof(double));
    if (st != IMP_OK) {
        imp_cnn_destroy(net);
        return st;
    }
    // He-normal weights on the trainable entries (fan-in 9 x real input
    // channels), zero biases; the output layer starts at zero so the
    // untrained network is the identity on its input.
    std::vector<float> h(NPARAM, 0.f);
    uint64_t s = seed;
    for (int layer = 0; layer < NLAYER - 1; ++layer) {
        const int fan_in = 9 * (layer == 0 ? CIO : CH);
        const double sd = std::sqrt(2.0 / fan_in);
        for (int r = 0; r < WSZ; ++r) {
            const int co = (r / CH) % CH, ci = r % CH;
            (void)co;
            if (layer == 0 && ci >= CIO) continue;
            const double u1 = unif(s), u2 = unif(s);
            h[(size_t)layer * PSZ + r] =
                (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2));
        }
    }
    if (hipMemcpy(net->params, h.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemset(net->adam_m, 0, NPARAM * sizeof(float)) != hipSuccess ||
        hipMemset(net->adam_v, 0, NPARAM * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        imp_cnn_destroy(net);
        return fail(IMP_EHIP, "imp_cnn_create: initialisation copy failed");
    }
    ctx->cnns.insert(net);
    *out = net;
    return IMP_OK;
}

extern "C" int imp_cnn_destroy(imp_cnn *net) {
    if (!net) return IMP_OK;
    // a network whose context was destroyed first has no device memory left
    if (net->ctx) {
        (void)hipStreamSynchronize(net->ctx->stream);
        imp_ctx *ctx = net->ctx;
        imp::detach_cnn(net);
        ctx->cnns.erase(net);
    }
    delete net;
    return IMP_OK;
}

namespace imp {
void detach_cnn(imp_cnn *net) {
    for (void *p : net->owned) (void)hipFree(p);
    net->owned.clear();
    net->ctx = nullptr;
}
}  // namespace imp

extern "C" int imp_cnn_num_params(void) { return NPARAM; }

extern "C" int imp_cnn_get_params(imp_cnn *net, float *host, int n) {
    if (!net || !net->ctx || !host || n != NPARAM)
        return fail(IMP_EINVAL, "imp_cnn_get_params: bad arguments or destroyed context");
    IMP_HIP(hipStreamSynchronize(net->ctx->stream));
    IMP_HIP(hipMemcpy(host, net->params, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    return IMP_OK;
}

extern "C" int imp_cnn_set_params(imp_cnn *net, const float *host, int n) {
    if (!net || !net->ctx || !host || n != NPARAM)
        return fail(IMP_EINVAL, "imp_cnn_set_params: bad arguments or destroyed context");
    for (int i = 0; i < NPARAM; ++i)
        if (!std::isfinite(host[i])) return fail(IMP_EINVAL, "imp_cnn_set_params: non-finite value");
    IMP_HIP(hipStreamSynchronize(net->ctx->stream));
    IMP_HIP(hipMemcpy(net->params, host, NPARAM * sizeof(float), hipMemcpyHostToDevice));
    IMP_HIP(hipMemset(net->adam_m, 0, NPARAM * sizeof(float)));
    IMP_HIP(hipMemset(net->adam_v, 0, NPARAM * sizeof(float)));
    net->step = 0;
    net->weights_dirty = true;
    return IMP_OK;
}

extern "C" int imp_cnn_forward(imp_cnn *net, const float *z_r, const float *z_i, float *x_r,
                               float *x_i) {
    if (!net || !net->ctx || !z_r || !z_i || !x_r || !x_i)
        return fail(IMP_EINVAL, "imp_cnn_forward: bad arguments");
    IMP_TRY(forward(net, z_r, z_i));
    hipLaunchKernelGGL(k_output, dim3(ceil_div(net->n, 256)), dim3(256), 0, net->ctx->stream, z_r,
                       z_i, net->net, net->n, net->dsc + RED_T, x_r, x_i);
    IMP_CHECK_LAUNCH();
    IMP_HIP(hipStreamSynchronize(net->ctx->stream));
    return IMP_OK;
}

extern "C" int imp_cnn_gradient(imp_cnn *net, const float *z_r, const float *z_i,
                                const float *t_r, const float *t_i, float *grad_host,
                                double *loss) {
    if (!net || !net->ctx || !z_r || !z_i || !t_r || !t_i)
        return fail(IMP_EINVAL, "imp_cnn_gradient: bad arguments");
    IMP_TRY(backward(net, z_r, z_i, t_r, t_i, loss));
    if (grad_host)
        IMP_HIP(hipMemcpy(grad_host, net->grad, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    return IMP_OK;
}

extern "C" int imp_cnn_train_step(imp_cnn *net, const float *z_r, const float *z_i,
                                  const float *t_r, const float *t_i, float lr, double *loss) {
    if (!net || !z_r || !z_i || !t_r || !t_i || !(lr > 0.f))
        return fail(IMP_EINVAL, "imp_cnn_train_step: bad arguments");
    IMP_TRY(backward(net, z_r, z_i, t_r, t_i, loss));
    net->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.f - std::pow(b1, (float)net->step);
    const float bc2 = 1.f - std::pow(b2, (float)net->step);
    hipLaunchKernelGGL(k_adam, dim3(ceil_div(NPARAM, 256)), dim3(256), 0, net->ctx->stream,
                       net->params, net->grad, net->adam_m, net->adam_v, lr, b1, b2, 1e-8f, bc1,
                       bc2);
    IMP_CHECK_LAUNCH();
    net->weights_dirty = true;
    return IMP_OK;
}
