// Internal declarations shared by the operator translation units.
#pragma once

#include "common.h"

#include <vector>

namespace imp {

// A real sparse matrix C (rows x cols) as two resident CSR copies, C and
// C^T, plus the work vector W C p [rows].
struct SparseReg {
    int rows = 0, cols = 0;
    long nnz = 0;
    int *c_ptr = nullptr, *c_idx = nullptr;
    float *c_val = nullptr;
    int *t_ptr = nullptr, *t_idx = nullptr;
    float *t_val = nullptr;
    float2 *cf = nullptr;
};

// Resident state of one Toeplitz normal operator A = sum_c S_c^H T S_c + lambda R.
struct ToepPlan {
    int N1 = 0, N2 = 0, N3 = 1;   // image [N1][N2][N3] = [Ny][Nx][Nz]
    int M1 = 0, M2 = 0, M3 = 1;   // 2x grid (M3 = 1 in 2D)
    long G = 0;                   // M1*M2*M3
    long nvox = 0;                // N1*N2*N3
    int nseg = 1;                 // L + 1 interpolators
    int ncoils = 1;
    int coils_per_group = 8;      // pass-E coil grouping
    int coils_per_group_min = 8;  // the grouping the partial buffer was sized for
    float tau = 0.f, t0 = 0.f;
    float lambda = 0.f;
    float2 *qf = nullptr;         // [nseg][G]  FFT(ifftshift(Q_l)), offsets -N zeroed
    float *qfr = nullptr;         // [nseg][G]  Re(qf), used when qf_real
    bool qf_real = false;         // qf is real to rounding (conjugate-symmetric Q_l)
    float2 *g1 = nullptr;         // [ncoils][N1][A][Ml] pass buffer
    float2 *partial = nullptr;    // [ngroups][nvox]
    float2 *sens = nullptr;       // [ncoils][nvox] or null (S = 1)
    float *fcyc = nullptr;        // [nvox] field map in Hz (w / 2pi)
    float *w = nullptr;           // [nvox] TV weight or null (W = I), not owned
    const SparseReg *sreg = nullptr;  // R = C^T W C instead of the stencil, not owned
    bool sreg_lost = false;           // sreg was destroyed while set: applies fail
    // CG work vectors (interleaved)
    float2 *vb = nullptr, *vx = nullptr, *vr = nullptr, *vp = nullptr,
           *vq = nullptr, *vbest = nullptr, *vz = nullptr;
    double *scal = nullptr;       // device CG scalars
    // circulant preconditioner M^-1 = F^H diag(pw) F on the image grid
    // (natural DFT order, 1/nvox of the inverse DFT folded in); all > 0
    float *pw = nullptr;
    float *pw_chan = nullptr;     // the Chan weights computed at create time
};

// The context's FFT twiddle table (64 KiB), made once at context creation.
int twiddle_table_create(imp_ctx *ctx);
int twiddles(imp_ctx *ctx, int L, const float2 **out);
int fft_grid(imp_ctx *ctx, float2 *data, int M1, int M2, int M3, bool inverse);
int toep_prepare_qf(imp_ctx *ctx, ToepPlan &P, const float *q_r, const float *q_i);
int toep_apply_normal(imp_ctx *ctx, ToepPlan &P, const float2 *p, float2 *out);
bool toep_sizes_supported(int N1, int N2, int N3);

// vector / stencil helpers (vec.hip)
int split_to_c(imp_ctx *ctx, const float *re, const float *im, float2 *out, long n);
int c_to_split(imp_ctx *ctx, const float2 *in, float *re, float *im, long n);
int fd_reg(imp_ctx *ctx, int N1, int N2, int N3, float lambda, const float2 *p,
           const float *w, float2 *out, bool accumulate);
int tv_weight(imp_ctx *ctx, int N1, int N2, int N3, const float2 *x, float *w);
// out (+)= lambda C^T W C p, W[r mod cols] on row r of C (sparse.hip)
int sparse_reg(imp_ctx *ctx, const SparseReg &R, float lambda, const float2 *p, const float *w,
               float2 *out, bool accumulate);
const SparseReg *sparse_reg_of(const imp_sparse_reg *reg);
imp_ctx *sparse_reg_ctx(const imp_sparse_reg *reg);
// Re(a^H b) in FP64, result to host (synchronous)
int dot_re(imp_ctx *ctx, const float2 *a, const float2 *b, long n, double *out);
int axpy_c(imp_ctx *ctx, float alpha, const float2 *x, float2 *y, long n);  // y += alpha x
int toep_cg(imp_ctx *ctx, ToepPlan &P, const float2 *b, float2 *x, int max_iter,
            double tol, int flags, imp_cg_stats *stats);
int toep_apply_full(imp_ctx *ctx, ToepPlan &P, const float2 *p, float2 *out);
// T. Chan's optimal circulant approximation of the summed Toeplitz kernel
// sum_l Q_l (standard Q grid) plus lambda times the periodic Laplacian;
// writes P.pw.
int precond_chan(imp_ctx *ctx, ToepPlan &P, const float *q_r, const float *q_i);
// z = M^-1 r
int precond_apply(imp_ctx *ctx, ToepPlan &P, const float2 *r, float2 *z);

// NUDFT (nudft.hip)
int nudft_adjoint(imp_ctx *ctx, int num_k, int num_x, int ncoils, const float *kx,
                  const float *ky, const float *kz, const float *t, const float *x,
                  const float *y, const float *z, const float *fm,
                  const float *phi_r, const float *phi_i, const float *d_r,
                  const float *d_i, const float *s_r, const float *s_i,
                  float *out_r, float *out_i);
int nudft_adjoint_noalloc(imp_ctx *ctx, int num_k, int num_x, int ncoils, const float *kx,
                          const float *ky, const float *kz, const float *t, const float *x,
                          const float *y, const float *z, const float *fm,
                          const float *phi_r, const float *phi_i, const float *d_r,
                          const float *d_i, const float *s_r, const float *s_i,
                          float *out_r, float *out_i);
int nudft_forward(imp_ctx *ctx, int num_k, int num_x, int ncoils, const float *kx,
                  const float *ky, const float *kz, const float *t, const float *x,
                  const float *y, const float *z, const float *fm,
                  const float *rho_r, const float *rho_i, const float *s_r,
                  const float *s_i, float *d_r, float *d_i);
int toeplitz_q(imp_ctx *ctx, int num_k, const float *kx, const float *ky,
               const float *kz, const float *t, const float *phi_r,
               const float *phi_i, int Nx, int Ny, int Nz, int num_q,
               const float *gx, const float *gy, const float *gz, int ntime_segs,
               int flags, const float *t_host_first_last, float *q_r, float *q_i);

// gridding path (grid.hip): Q_l on the standard 2x grid, F^H d on [Ny][Nx][Nz]
int gridding_q(imp_ctx *ctx, int num_k, const float *kx, const float *ky, const float *kz,
               const float *t, int Nx, int Ny, int Nz, int ntime_segs, float gridOS, float *q_r,
               float *q_i);
int gridding_fhd(imp_ctx *ctx, int num_k, int ncoils, const float *kx, const float *ky,
                 const float *kz, const float *t, int Nx, int Ny, int Nz, const float *fm,
                 const float *d_r, const float *d_i, const float *s_r, const float *s_i,
                 int ntime_segs, float gridOS, float *out_r, float *out_i);

int bruteforce_cg(imp_ctx *ctx, int num_k, int num_x, int ncoils, const float *kx,
                  const float *ky, const float *t, const float *x, const float *y,
                  const float *fm, const float *d_r, const float *d_i, const float *s_r,
                  const float *s_i, int cg_num, float lambda, const SparseReg *sreg,
                  int tv_num, float *x_r, float *x_i, imp_tv_pass_hook hook = nullptr,
                  void *hook_user = nullptr);

// internal CG flag: track the best iterate in ToepPlan::vbest but leave the
// last iterate in x
constexpr int CG_KEEP_LAST = 0x100;
// internal CG flag (plain CG only): start with the residual norm and best
// residual the previous solve on this plan ended with, instead of |b - A x0|^2
// (the restarts of toeplitz_recon, recon.cu:2043-2048)
constexpr int CG_CARRY_SCALARS = 0x200;

}  // namespace imp

struct imp_toep {
    imp_ctx *ctx = nullptr;
    imp::ToepPlan P;
    std::vector<void *> owned;
};
