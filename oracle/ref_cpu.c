/*
 * C restatement of the reference CPU path -- TEST / BASELINE INFRASTRUCTURE
 * ONLY.  Used by tests/ as an independent FP32 implementation and by
 * bench.py's cpu_baseline leg as the timed reference CPU path.  The product
 * library never links or calls it.
 *
 * Follows kernels/ftCpu.cpp:110-277 (ftCpu / iftCpu): the same loop order
 * (samples outer for ft, voxels outer for ift), FP32 arithmetic, FP32
 * sinf/cosf, and OpenMP on the inner reduction (ftCpu.cpp:162-167,
 * 248-253).  The SENSE wrap follows kernels/parImagingCpu.cpp:85-154, the
 * direct F^H d of toeplitz/computeFH.cmem.cu:155-200 (3D phase, conj(phi) d)
 * and the direct Q of toeplitz/computeQ.cmem.cu:169-208 are restated in the
 * same FP32 style (the reference has no CPU version of those two).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define TPI 6.28318530717958647692f

/* kdata = F idata (one coil): ftCpu, kernels/ftCpu.cpp:110-189 */
void ft_cpu(float *kdata_r, float *kdata_i, const float *idata_r, const float *idata_i,
            const float *kx, const float *ky, const float *ix, const float *iy,
            const float *fm, const float *t, int num_k, int num_i) {
    for (int i = 0; i < num_k; i++) {
        float sumr = 0.0f, sumi = 0.0f;
        const float kxtpi = kx[i] * TPI, kytpi = ky[i] * TPI, ti = t[i];
#pragma omp parallel for reduction(+ : sumr, sumi)
        for (int j = 0; j < num_i; j++) {
            const float expr = (kxtpi * ix[j] + kytpi * iy[j]) + (fm[j] * ti);
            const float c = cosf(expr), s = sinf(expr);
            sumr += (c * idata_r[j]) + (s * idata_i[j]);
            sumi += (-s * idata_r[j]) + (c * idata_i[j]);
        }
        kdata_r[i] = sumr;
        kdata_i[i] = sumi;
    }
}

/* idata = F^H kdata (one coil): iftCpu, kernels/ftCpu.cpp:199-277 */
void ift_cpu(float *idata_r, float *idata_i, const float *kdata_r, const float *kdata_i,
             const float *kx, const float *ky, const float *ix, const float *iy,
             const float *fm, const float *t, int num_k, int num_i) {
    for (int j = 0; j < num_i; j++) {
        float sumr = 0.0f, sumi = 0.0f;
        const float xt = ix[j] * TPI, yt = iy[j] * TPI, fj = fm[j];
#pragma omp parallel for reduction(+ : sumr, sumi)
        for (int i = 0; i < num_k; i++) {
            const float expr = (kx[i] * xt + ky[i] * yt) + (fj * t[i]);
            const float c = cosf(expr), s = sinf(expr);
            sumr += (c * kdata_r[i]) - (s * kdata_i[i]);
            sumi += (s * kdata_r[i]) + (c * kdata_i[i]);
        }
        idata_r[j] = sumr;
        idata_i[j] = sumi;
    }
}

/* parallelIftCpu (kernels/parImagingCpu.cpp:114-154): sum_c conj(S_c) F^H d_c */
void parallel_ift_cpu(float *out_r, float *out_i, const float *kdata_r, const float *kdata_i,
                      const float *kx, const float *ky, const float *ix, const float *iy,
                      const float *fm, const float *t, int num_k, int num_i,
                      const float *s_r, const float *s_i, int ncoils) {
    float *tr = (float *)malloc(sizeof(float) * num_i);
    float *ti = (float *)malloc(sizeof(float) * num_i);
    memset(out_r, 0, sizeof(float) * num_i);
    memset(out_i, 0, sizeof(float) * num_i);
    for (int c = 0; c < ncoils; c++) {
        ift_cpu(tr, ti, kdata_r + (size_t)c * num_k, kdata_i + (size_t)c * num_k, kx, ky, ix, iy,
                fm, t, num_k, num_i);
        for (int j = 0; j < num_i; j++) {
            const float sr = s_r ? s_r[(size_t)c * num_i + j] : 1.0f;
            const float si = s_i ? s_i[(size_t)c * num_i + j] : 0.0f;
            out_r[j] += sr * tr[j] + si * ti[j];
            out_i[j] += sr * ti[j] - si * tr[j];
        }
    }
    free(tr);
    free(ti);
}

/* parallelFtCpu (kernels/parImagingCpu.cpp:85-112): d_c = F (S_c x) */
void parallel_ft_cpu(float *kdata_r, float *kdata_i, const float *x_r, const float *x_i,
                     const float *kx, const float *ky, const float *ix, const float *iy,
                     const float *fm, const float *t, int num_k, int num_i,
                     const float *s_r, const float *s_i, int ncoils) {
    float *vr = (float *)malloc(sizeof(float) * num_i);
    float *vi = (float *)malloc(sizeof(float) * num_i);
    for (int c = 0; c < ncoils; c++) {
        for (int j = 0; j < num_i; j++) {
            const float sr = s_r ? s_r[(size_t)c * num_i + j] : 1.0f;
            const float si = s_i ? s_i[(size_t)c * num_i + j] : 0.0f;
            vr[j] = sr * x_r[j] - si * x_i[j];
            vi[j] = sr * x_i[j] + si * x_r[j];
        }
        ft_cpu(kdata_r + (size_t)c * num_k, kdata_i + (size_t)c * num_k, vr, vi, kx, ky, ix, iy,
               fm, t, num_k, num_i);
    }
    free(vr);
    free(vi);
}

/* Direct F^H d with 3D phase and conj(phi) d weighting, one coil, FP32:
 * the arithmetic of ComputeFH_GPU_BF (toeplitz/computeFH.cmem.cu:155-200)
 * in the loop order of iftCpu. */
void fhd_cpu(float *out_r, float *out_i, const float *rhophi_r, const float *rhophi_i,
             const float *kx, const float *ky, const float *kz, const float *t, const float *ix,
             const float *iy, const float *iz, const float *fm, int num_k, int num_i) {
    for (int j = 0; j < num_i; j++) {
        float sr = 0.0f, si = 0.0f;
        const float xj = ix[j], yj = iy[j], zj = iz ? iz[j] : 0.0f, fj = fm[j];
#pragma omp parallel for reduction(+ : sr, si)
        for (int k = 0; k < num_k; k++) {
            const float arg = TPI * kx[k] * xj + TPI * ky[k] * yj +
                              (kz ? TPI * kz[k] * zj : 0.0f) + t[k] * fj;
            const float c = cosf(arg), s = sinf(arg);
            sr += rhophi_r[k] * c - rhophi_i[k] * s;
            si += rhophi_i[k] * c + rhophi_r[k] * s;
        }
        out_r[j] = sr;
        out_i[j] = si;
    }
}

/* Direct Q_l for one l, FP32: ComputeQ_GPU (toeplitz/computeQ.cmem.cu:169-208)
 * with hanning_d (:125-138). */
void q_cpu(float *q_r, float *q_i, const float *kx, const float *ky, const float *kz,
           const float *t, const float *phimag, const float *gx, const float *gy,
           const float *gz, int num_k, int num_q, float l, float tau, float t0) {
    for (int j = 0; j < num_q; j++) {
        float sr = 0.0f, si = 0.0f;
        const float xj = gx[j], yj = gy[j], zj = gz ? gz[j] : 0.0f;
#pragma omp parallel for reduction(+ : sr, si)
        for (int k = 0; k < num_k; k++) {
            const float a = t[k] - tau * l - t0;
            float atm = 0.0f;
            if (fabsf(a) < tau) atm = 0.5f + 0.5f * cosf(3.14159265358979f * a / tau);
            const float arg = TPI * kx[k] * xj + TPI * ky[k] * yj + (kz ? TPI * kz[k] * zj : 0.0f);
            const float w = atm * (phimag ? phimag[k] : 1.0f);
            sr += w * cosf(arg);
            si += w * sinf(arg);
        }
        q_r[j] = sr;
        q_i[j] = si;
    }
}

/* lam * sum_d D_d^T D_d p on a row-major [ny][nx] image with periodic first
 * differences (the D of createDWithPeriodicBoundary.m, W = 1): the operator
 * of DHWD2dCpu (kernels/DHWD2dCpu.cpp:96-325) in this image layout. */
static void fd_cpu(float *s_r, float *s_i, const float *p_r, const float *p_i, int nx, int ny,
                   float lam) {
    for (int y = 0; y < ny; y++) {
        for (int x = 0; x < nx; x++) {
            const int i = y * nx + x;
            const int xp = y * nx + (x + 1) % nx, xm = y * nx + (x + nx - 1) % nx;
            const int yp = ((y + 1) % ny) * nx + x, ym = ((y + ny - 1) % ny) * nx + x;
            s_r[i] = lam * (4.0f * p_r[i] - p_r[xp] - p_r[xm] - p_r[yp] - p_r[ym]);
            s_i[i] = lam * (4.0f * p_i[i] - p_i[xp] - p_i[xm] - p_i[yp] - p_i[ym]);
        }
    }
}

static float dot_cpu(const float *a_r, const float *a_i, const float *b_r, const float *b_i,
                     int n) {
    /* dotProductCpu (kernels/dotProductCpu.cpp:84-113): real part, FP32 */
    float s = 0.0f;
    for (int i = 0; i < n; i++) s += a_r[i] * b_r[i] + a_i[i] * b_i[i];
    return s;
}

/* Exported for the timed CPU baseline of the vector stages (one thread, as
 * the reference's DHWD2dCpu / dotProductCpu / addCpu run): the stencil, the
 * real dot product and y += a x (addCpu.cpp) on split re/im arrays. */
void fd_reg_cpu(float *s_r, float *s_i, const float *p_r, const float *p_i, int nx, int ny,
                float lam) {
    fd_cpu(s_r, s_i, p_r, p_i, nx, ny, lam);
}

float dot_re_cpu(const float *a_r, const float *a_i, const float *b_r, const float *b_i, int n) {
    return dot_cpu(a_r, a_i, b_r, b_i, n);
}

void axpy_cpu(float *y_r, float *y_i, float a, const float *x_r, const float *x_i, int n) {
    for (int i = 0; i < n; i++) {
        y_r[i] += a * x_r[i];
        y_i[i] += a * x_i[i];
    }
}

/* bruteForceCpu's CG (bruteForce/bruteForceCpu.cpp:121-534, the loop at
 * :280-420) with the finite-difference regulariser: cg_num - 1 iterations
 * from x (in/out), grad = F^H (y - A f) - lam D^T D x, d = grad + gamma d,
 * alpha = Re(d^H grad) / (|F d|^2 + Re((lam D^T D d)^H d)), A f += alpha F d,
 * x += alpha d.  FP32 vectors and scalars (FLOAT_T = float). */
void bruteforce_cpu(float *x_r, float *x_i, const float *y_r, const float *y_i,
                    const float *kx, const float *ky, const float *ix, const float *iy,
                    const float *fm, const float *t, const float *s_r, const float *s_i,
                    int ncoils, int num_k, int nx, int ny, int cg_num, float lam) {
    const int ni = nx * ny, nk = num_k * ncoils;
    float *af_r = calloc(nk, sizeof(float)), *af_i = calloc(nk, sizeof(float));
    float *q_r = calloc(nk, sizeof(float)), *q_i = calloc(nk, sizeof(float));
    float *r_r = calloc(nk, sizeof(float)), *r_i = calloc(nk, sizeof(float));
    float *g_r = calloc(ni, sizeof(float)), *g_i = calloc(ni, sizeof(float));
    float *d_r = calloc(ni, sizeof(float)), *d_i = calloc(ni, sizeof(float));
    float *c_r = calloc(ni, sizeof(float)), *c_i = calloc(ni, sizeof(float));
    parallel_ft_cpu(af_r, af_i, x_r, x_i, kx, ky, ix, iy, fm, t, num_k, ni, s_r, s_i, ncoils);
    float oldinprod = 0.0f;
    for (int it = 1; it < cg_num; it++) {
        for (int i = 0; i < nk; i++) {
            r_r[i] = y_r[i] - af_r[i];
            r_i[i] = y_i[i] - af_i[i];
        }
        parallel_ift_cpu(g_r, g_i, r_r, r_i, kx, ky, ix, iy, fm, t, num_k, ni, s_r, s_i, ncoils);
        if (lam != 0.0f) {
            fd_cpu(c_r, c_i, x_r, x_i, nx, ny, lam);
            for (int i = 0; i < ni; i++) {
                g_r[i] -= c_r[i];
                g_i[i] -= c_i[i];
            }
        }
        const float newinprod = dot_cpu(g_r, g_i, g_r, g_i, ni);
        if (it == 1) {
            memcpy(d_r, g_r, ni * sizeof(float));
            memcpy(d_i, g_i, ni * sizeof(float));
        } else {
            const float gamma = oldinprod == 0.0f ? 0.0f : newinprod / oldinprod;
            for (int i = 0; i < ni; i++) {
                d_r[i] = g_r[i] + gamma * d_r[i];
                d_i[i] = g_i[i] + gamma * d_i[i];
            }
        }
        oldinprod = newinprod;
        parallel_ft_cpu(q_r, q_i, d_r, d_i, kx, ky, ix, iy, fm, t, num_k, ni, s_r, s_i, ncoils);
        float a_ctcd = 0.0f;
        if (lam != 0.0f) {
            fd_cpu(c_r, c_i, d_r, d_i, nx, ny, lam);
            a_ctcd = dot_cpu(c_r, c_i, d_r, d_i, ni);
        }
        const float alpha = dot_cpu(d_r, d_i, g_r, g_i, ni) / (dot_cpu(q_r, q_i, q_r, q_i, nk) + a_ctcd);
        for (int i = 0; i < nk; i++) {
            af_r[i] += alpha * q_r[i];
            af_i[i] += alpha * q_i[i];
        }
        for (int i = 0; i < ni; i++) {
            x_r[i] += alpha * d_r[i];
            x_i[i] += alpha * d_i[i];
        }
    }
    free(af_r); free(af_i); free(q_r); free(q_i); free(r_r); free(r_i);
    free(g_r); free(g_i); free(d_r); free(d_i); free(c_r); free(c_i);
}

/* ---------------------------------------------------------------------------
 * Gold CPU gridding: gridding_Gold_2D / gridding_Gold_3D
 * (toeplitz/CPU_kernels.cpp:89-330), the reference's CPU version of the
 * gridding step of computeQ_GPU_Grid / computeFH_GPU_Grid.  FP32 throughout,
 * the fixed beta = 18.5547 of the gold path (CPU_kernels.cpp:110), I0 by the
 * polynomial approximations of Abramowitz & Stegun 9.8.1-9.8.2, samples
 * scaled by `kscale` (2 for Q, 1 for F^H d) and weighted by a_l(t) of
 * segment l, window indices clamped to [0, gridOS n - 1] (no wrap).
 * Output in the reference's order after its reorder: 2D [x][y] (y fastest),
 * 3D [y][x][z] (z fastest); grid sizes g[3] (g[2] = 1 in 2D).  Data re/im of
 * NULL means unit data (the Q case).  Single-threaded, as the reference. */
static float gold_i0(float x) {
    float ax = fabsf(x), y;
    if (ax < 3.75f) {
        y = x / 3.75f;
        y *= y;
        return 1.0f + y * (3.5156229f + y * (3.0899424f + y * (1.2067492f + y * (0.2659732f +
                     y * (0.0360768f + y * 0.0045813f)))));
    }
    y = 3.75f / ax;
    return (expf(ax) / sqrtf(ax)) * (0.39894228f + y * (0.01328592f + y * (0.00225319f +
           y * (-0.00157565f + y * (0.00916281f + y * (-0.02057706f + y * (0.02635537f +
           y * (-0.01647633f + y * 0.00392377f))))))));
}

static float gold_hanning(float tm, float tau, float l, float t0) {
    const float x = tm - tau * l - t0;
    return fabsf(x) < tau ? 0.5f + 0.5f * cosf(3.14159265359f * x / tau) : 0.0f;
}

static float gold_kb(float s, int g, float os, float width, float beta) {
    const float dist = fabsf(s - (float)g) / os;
    const float u = 2.0f * dist / width;
    const float v = gold_i0(beta * sqrtf(1.0f - u * u)) / width;
    return v != v ? 0.0f : v;   /* NaN -> 0 as CPU_kernels.cpp:162-163 */
}

void gridding_gold(int m, const float *kx, const float *ky, const float *kz, float kscale,
                   const float *re, const float *im, const float *t, float l, float tau,
                   int nx, int ny, int nz, float gridOS, float beta, const int *g,
                   float *grid_r, float *grid_i) {
    const float width = 4.0f, half = width * gridOS / 2.0f;
    const int d3 = nz > 1;
    const long ng = (long)g[0] * g[1] * g[2];
    float *ar = (float *)calloc(ng, sizeof(float)), *ai = (float *)calloc(ng, sizeof(float));
    const float t0 = t[0];
    for (int i = 0; i < m; i++) {
        const float atm = gold_hanning(t[i], tau, l, t0);
        const float vr = re ? re[i] : 1.0f, vi = re ? im[i] : 0.0f;
        const float sx = gridOS * (kscale * kx[i] + (float)nx / 2.0f);
        const float sy = gridOS * (kscale * ky[i] + (float)ny / 2.0f);
        const float sz = d3 ? gridOS * (kscale * kz[i] + (float)nz / 2.0f) : 0.0f;
        const int xl = (int)fmaxf(0.0f, ceilf(sx - half));
        const int xh = (int)fminf(gridOS * (float)nx - 1.0f, floorf(sx + half));
        const int yl = (int)fmaxf(0.0f, ceilf(sy - half));
        const int yh = (int)fminf(gridOS * (float)ny - 1.0f, floorf(sy + half));
        const int zl = d3 ? (int)fmaxf(0.0f, ceilf(sz - half)) : 0;
        const int zh = d3 ? (int)fminf(gridOS * (float)nz - 1.0f, floorf(sz + half)) : 0;
        for (int z = zl; z <= zh; z++) {
            const float kbz = d3 ? gold_kb(sz, z, gridOS, width, beta) : 1.0f;
            for (int y = yl; y <= yh; y++) {
                const float kby = gold_kb(sy, y, gridOS, width, beta);
                for (int x = xl; x <= xh; x++) {
                    const float w = gold_kb(sx, x, gridOS, width, beta) * kby * kbz;
                    const long idx = x + (long)y * g[0] + (long)z * g[0] * g[1];
                    ar[idx] += w * vr * atm;
                    ai[idx] += w * vi * atm;
                }
            }
        }
    }
    /* reorder x->y->z to the reference's output order (CPU_kernels.cpp:183-197, 310-325) */
    for (int x = 0; x < g[0]; x++)
        for (int y = 0; y < g[1]; y++)
            for (int z = 0; z < g[2]; z++) {
                const long src = x + (long)y * g[0] + (long)z * g[0] * g[1];
                const long dst = d3 ? z + (long)x * g[2] + (long)y * g[0] * g[2]
                                    : y + (long)x * g[0];
                grid_r[dst] = ar[src];
                grid_i[dst] = ai[src];
            }
    free(ar);
    free(ai);
}
