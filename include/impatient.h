/*
 * impatient.h -- C ABI of the MI355X (gfx950) IMPATIENT reconstruction engine.
 *
 * Every entry point returns an int status (IMP_OK == 0, negative on error);
 * imp_last_error() returns the message of the calling thread's last failure.
 * No entry point calls exit(): errors the reference handles with
 * printf + exit(1) come back as status codes here.
 *
 * Memory: functions named imp_<op> take DEVICE pointers (HBM-resident
 * operands, asynchronous on the context's HIP stream). The file-level
 * wrappers (imp_toeplitz_computeQ_GPU, imp_toeplitz_computeFH_GPU,
 * imp_toeplitz_recon, imp_bruteForceGpu) take HOST pointers exactly like the
 * reference functions they replace and are synchronous.
 *
 * Complex vectors are passed as split real / imaginary float arrays, the
 * layout of the reference's data files and host arrays (kdata_r/kdata_i,
 * sensi_r/sensi_i, Fhd_R/Fhd_I, Qr[l]/Qi[l]).
 *
 * Phase convention (reference README "Data Format", kernels/ftGpu.cu:493):
 *   forward  d_c(k) = sum_x S_c(x) rho(x) exp(-i(2*pi*k.x + w(x) t_k))
 *   adjoint  rho(x) = sum_c conj(S_c(x)) sum_k d_c(k) exp(+i(2*pi*k.x + w(x) t_k))
 * k in cycles/FOV, x in FOV units, w (field map) in rad/s, t in seconds.
 */
#ifndef IMPATIENT_H
#define IMPATIENT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define IMP_OK            0
#define IMP_EINVAL       -1   /* bad argument / size                          */
#define IMP_EHIP         -2   /* HIP runtime error                            */
#define IMP_EIO          -3   /* file missing / short read / write failure    */
#define IMP_ENOMEM       -4   /* allocation failed                            */
#define IMP_EFORMAT      -5   /* malformed data file header                   */
#define IMP_ENODEV       -6   /* no usable GPU                                */
#define IMP_EBREAKDOWN   -7   /* CG breakdown: p^H A p zero or not finite      */
#define IMP_EUNSUPPORTED -8   /* size the kernels do not support (yet)        */

const char *imp_version_string(void);
const char *imp_status_string(int status);
/* Message of the calling thread's last failed call ("" if none). */
const char *imp_last_error(void);

/* Process-wide reference-compatibility switches; 0 (default) = correct math.
 * IMP_COMPAT_Q_CUMULATIVE: imp_toeplitz_computeQ_GPU returns the prefix sums
 * over l that the reference direct path produces (its device accumulator is
 * not cleared between segments, computeQ.cmem.cu:190-191, 951-963). */
#define IMP_COMPAT_Q_CUMULATIVE 0x1
int imp_set_compat(int flags);
int imp_get_compat(void);

/* --------------------------------------------------------------- context */
typedef struct imp_ctx imp_ctx;

/* One context per (host thread, GPU); owns one HIP stream and a workspace.
 * Replaces the cudaSetDevice + implicit default stream of main_mri.cu:222-233. */
int imp_ctx_create(int device, imp_ctx **out);
int imp_ctx_destroy(imp_ctx *ctx);
/* hipStream_t of the context (as void*). */
int imp_ctx_stream(imp_ctx *ctx, void **hip_stream);
int imp_ctx_synchronize(imp_ctx *ctx);

/* Kernel timer: when enabled, the context records HIP events on its stream
 * around its main kernels.  imp_profile_read sums the recorded durations of
 * one kernel name ("k_nudft_mfma", "k_nudft_valu", "toeplitz_apply", ...)
 * since the last reset.  Stands in for the reference's mriTimer singleton
 * (xcpplib timers in main_mri.cu / recon.cu), per context instead of global. */
int imp_profile_enable(imp_ctx *ctx, int on);
int imp_profile_reset(imp_ctx *ctx);
int imp_profile_read(imp_ctx *ctx, const char *kernel, double *total_ms, int *launches);

/* Device-memory helpers for callers without an allocator of their own. */
int imp_device_alloc(imp_ctx *ctx, size_t bytes, void **dptr);
int imp_device_free(imp_ctx *ctx, void *dptr);
int imp_copy_h2d(imp_ctx *ctx, void *dst, const void *src, size_t bytes);
int imp_copy_d2h(imp_ctx *ctx, void *dst, const void *src, size_t bytes);
int imp_memset_d(imp_ctx *ctx, void *dst, int value, size_t bytes);
/* Page-locked host memory (hipHostMalloc) for staging uploads: copies from
 * it run at the link's DMA rate.  No counterpart in the reference; for
 * callers that upload large inputs repeatedly (the batched F^H d). */
int imp_host_alloc(imp_ctx *ctx, size_t bytes, void **hptr);
int imp_host_free(imp_ctx *ctx, void *hptr);

/* ------------------------------------------------------------- operators
 * Direct, field-corrected NUDFT with SENSE (device pointers).
 *
 * imp_nudft_adjoint replaces ComputeFH_GPU_BF / computeFH_GPU_BF /
 * parcomputeAH_GPU_BF (toeplitz/computeFH.cmem.cu:155-200, 805-869,
 * 1146-1186) and iftGpu + parallelIftGpu (kernels/ftGpu.cu:653-1013,
 * kernels/parImagingGpu.cu:119-152):
 *   out(x) = sum_c conj(S_c(x)) sum_k conj(phi_c(k)) d_c(k) e^{+i(2pi k.x + fm(x) t_k)}
 *   kx,ky,t [num_k]; kz [num_k] or NULL (2D); x,y [num_x]; z NULL (2D);
 *   fm [num_x] or NULL (0); phi_r/phi_i [ncoils*num_k] or NULL (phi = 1);
 *   d_r/d_i [ncoils*num_k] coil-major; s_r/s_i [ncoils*num_x] or NULL (S = 1);
 *   out_r/out_i [num_x].
 */
int imp_nudft_adjoint(imp_ctx *ctx, int num_k, int num_x, int ncoils,
                      const float *kx, const float *ky, const float *kz,
                      const float *t,
                      const float *x, const float *y, const float *z,
                      const float *fm,
                      const float *phi_r, const float *phi_i,
                      const float *d_r, const float *d_i,
                      const float *s_r, const float *s_i,
                      float *out_r, float *out_i);

/* imp_nudft_adjoint without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and result as imp_nudft_adjoint (stacked 3D trajectories are
 * summed directly, without the separable form).  Every buffer it needs comes
 * from one fixed 64 MiB device array per GPU, held until the process ends;
 * a call that needs more returns IMP_EUNSUPPORTED before launching anything
 * (about 2D 256^2 voxels with 8 coils and 32768 samples fits).  All calls on
 * one GPU share that array: calls that overlap on one GPU (two streams, two
 * contexts, two threads) race and must be serialised by the caller. */
int imp_nudft_adjoint_noalloc(imp_ctx *ctx, int num_k, int num_x, int ncoils,
                              const float *kx, const float *ky, const float *kz,
                              const float *t,
                              const float *x, const float *y, const float *z,
                              const float *fm,
                              const float *phi_r, const float *phi_i,
                              const float *d_r, const float *d_i,
                              const float *s_r, const float *s_i,
                              float *out_r, float *out_i);

/* Replaces ftGpu + parallelFtGpu (kernels/ftGpu.cu:198-604,
 * kernels/parImagingGpu.cu:92-117):
 *   d_c(k) = sum_x S_c(x) rho(x) e^{-i(2pi k.x + fm(x) t_k)}
 *   d_r/d_i [ncoils*num_k] coil-major.  No power-of-two padding is needed. */
int imp_nudft_forward(imp_ctx *ctx, int num_k, int num_x, int ncoils,
                      const float *kx, const float *ky, const float *kz,
                      const float *t,
                      const float *x, const float *y, const float *z,
                      const float *fm,
                      const float *rho_r, const float *rho_i,
                      const float *s_r, const float *s_i,
                      float *d_r, float *d_i);

/* Batched F^H d: several images (different sizes, trajectories, coil maps;
 * each 3..32 coils, all 2D or all 3D) in one persistent MFMA launch whose
 * flattened grid covers every image's point tiles.  Arguments of each job
 * as for imp_nudft_adjoint.  chunk_scale: workgroups per CU targeted per job
 * (the schedule knob an autotuner sweeps; 0 = 8).  No counterpart in the
 * reference, which loops slices on the host (main_mri.cu:771-904). */
typedef struct {
    int num_k, num_x, ncoils;
    const float *kx, *ky, *kz, *t, *x, *y, *z, *fm;
    const float *phi_r, *phi_i, *d_r, *d_i, *s_r, *s_i;
    float *out_r, *out_i;
} imp_nudft_job;
int imp_nudft_adjoint_batch(imp_ctx *ctx, int njobs, const imp_nudft_job *jobs,
                            int chunk_scale);

/* Toeplitz kernels Q_l, l = 0..ntime_segs (ComputeQ_GPU / computeQ_GPU /
 * toeplitz_computeQ_GPU, toeplitz/computeQ.cmem.cu:169-208, 756-822, 826-1028):
 *   Q_l(r) = sum_k a_l(t_k) |phi(k)|^2 e^{+i 2pi k.r}
 *   a_l(t) = 1/2 + 1/2 cos(pi (t - t0 - l tau)/tau) if |t - t0 - l tau| < tau, else 0
 *   tau = (t[num_k-1] - t[0]) / ntime_segs (ntime_segs > 0), t0 = t[0].
 * gx/gy/gz: coordinates of the num_q grid points (the ixQ/iyQ/izQ files of
 * main_mri.cu:461-547), or all NULL for the standard grid of (Nx,Ny,Nz):
 * index ((y+Ny)*2Nx + (x+Nx))*M3 + (z+Nz), r = (x/Nx, y/Ny, z/Nz), M3 = 2Nz
 * (3D) or 1 (2D), num_q = 4NxNy*M3.  gz may be NULL in 2D.
 * phi_r/phi_i [num_k] or NULL (phi = 1).  q_r/q_i [(ntime_segs+1)*num_q].
 * flags: IMP_Q_CUMULATIVE reproduces the reference direct path, whose device
 * accumulator is not cleared between segments (computeQ.cmem.cu:190-191,
 * 951-963), i.e. returns prefix sums over l.  On the standard grid the
 * sums are evaluated as a separable GEMM (e^{i 2pi k.r} = e_x e_y e_z per
 * sample, each segment over the samples where a_l != 0); IMP_Q_DIRECT
 * evaluates every grid point with the per-point kernel instead. */
#define IMP_Q_CUMULATIVE 0x1
#define IMP_Q_DIRECT 0x2
int imp_toeplitz_q(imp_ctx *ctx, int num_k,
                   const float *kx, const float *ky, const float *kz,
                   const float *t, const float *phi_r, const float *phi_i,
                   int Nx, int Ny, int Nz, int num_q,
                   const float *gx, const float *gy, const float *gz,
                   int ntime_segs, int flags,
                   float *q_r, float *q_i);

/* Q_l by Kaiser-Bessel gridding (the -toeplitzGridding path:
 * computeQ_GPU_Grid, toeplitz/computeQ.cmem.cu:401-754, called per l from
 * toeplitz_computeQ_GPU): the samples at 2k, weights a_l(t) (phi = 1), on
 * the 2N grid of the standard Q layout (as imp_toeplitz_q with NULL
 * coordinates), q_r/q_i [(ntime_segs+1) * 4NxNy*M3].  Kernel width 4,
 * Beatty's beta, oversampling gridOS rounded as the reference rounds it (x/y
 * grid even in 2D, z grid a multiple of 4 in 3D), no wrap at the grid edge,
 * deapodization by the kernel's transform.  kz may be NULL in 2D. */
int imp_toeplitz_q_gridding(imp_ctx *ctx, int num_k, const float *kx, const float *ky,
                            const float *kz, const float *t, int Nx, int Ny, int Nz,
                            int ntime_segs, float gridOS, float *q_r, float *q_i);

/* F^H d by gridding with time segmentation (parcomputeAH_GPU_Grid /
 * computeAH_GPU_Grid / computeFH_GPU_Grid, toeplitz/computeFH.cmem.cu:470-802,
 * 1032-1117):
 *   out = sum_c conj(S_c) sum_l e^{+i fm (t0 + l tau)} grid(a_l(t) d_c)
 * on the image grid x = (i - N/2)/N, out [Ny][Nx][Nz] (x fastest in 2D, z in
 * 3D); d_r/d_i [ncoils*num_k] coil-major; s_r/s_i [ncoils*Nx*Ny*Nz] or
 * NULL (S = 1); fm [Nx*Ny*Nz] or NULL (0). */
int imp_fhd_gridding(imp_ctx *ctx, int num_k, int ncoils, const float *kx, const float *ky,
                     const float *kz, const float *t, int Nx, int Ny, int Nz, const float *fm,
                     const float *d_r, const float *d_i, const float *s_r, const float *s_i,
                     int ntime_segs, float gridOS, float *out_r, float *out_i);

/* Periodic first-difference regularizer (DHWD2dGpu / DHWD3dGpu,
 * toeplitz/DHWD2dGpu_toeplitz.cu:132-387, toeplitz/DHWD3dGpu_toeplitz.cu:170-580):
 *   s = lambda * sum_d D_d^T W D_d p,  (D_d p)(i) = p(i) - p(i + e_d mod N_d)
 * image [N1][N2][N3] row-major (N3 = 1 in 2D); w [N1*N2*N3] or NULL (W = I).
 * Because the operator sums over all axes, the column-major brute-force
 * variant (kernels/DHWD2dGpu.cu) gives identical results on square images. */
int imp_fd_reg_apply(imp_ctx *ctx, int N1, int N2, int N3, float lambda,
                     const float *p_r, const float *p_i, const float *w,
                     float *s_r, float *s_i);

/* TV weight w = 0.5 / sqrt(sum_d |D_d x|^2 + 1e-6)
 * (multiply2dGpu / multiply3dGpu, toeplitz/multiplyGpu_toeplitz.cu:81-172). */
int imp_tv_weight(imp_ctx *ctx, int N1, int N2, int N3,
                  const float *x_r, const float *x_i, float *w);

/* ------------------------------------------------ Toeplitz normal operator
 * A p = sum_c conj(S_c) sum_l e^{+i w tau_l} T_l (e^{-i w tau_l} S_c p) + lambda R p
 * with tau_l = t0 + l*tau and T_l the 2x-oversampled circulant embedding of
 * Q_l (CudaFhF / CudaAhA / parCudaAhA / Computeq, toeplitz/recon.cu:1113-1697).
 * Image layout [N1][N2][N3] = [Ny][Nx][Nz] row-major, N3 = 1 for 2D.
 * q_r/q_i: DEVICE pointers to Q_l on the standard grid (imp_toeplitz_q with
 * NULL coordinates, without IMP_Q_CUMULATIVE), (ntime_segs+1) planes.
 * The plan keeps FFT(ifftshift(Q_l)) resident in HBM (the reference
 * re-uploads it from the host for every coil and segment, recon.cu:1404-1411).
 * fm [N], s_r/s_i [ncoils*N] (NULL: single coil with S = 1) are copied. */
typedef struct imp_toep imp_toep;
int imp_toep_create(imp_ctx *ctx, int N1, int N2, int N3, int ncoils,
                    int ntime_segs, float tau, float t0,
                    const float *q_r, const float *q_i,
                    const float *fm, const float *s_r, const float *s_i,
                    float lambda, imp_toep **out);
int imp_toep_destroy(imp_toep *op);
/* Replace the TV weight W of lambda R (device [N], NULL = identity). */
int imp_toep_set_weight(imp_toep *op, const float *w);
/* Use lambda C^T W C (a sparse-matrix regularizer, below) as R instead of
 * the finite-difference stencil: the -reg branch of Computeq
 * (toeplitz/recon.cu:1543-1640).  C must be (ndim N) x N, ndim = 2 for
 * N3 = 1 and 3 otherwise, and created on the operator's context
 * (IMP_EINVAL otherwise); reg is not copied.  Destroying reg while it is
 * set makes the operator's applies fail (IMP_EINVAL) until it is set again;
 * NULL restores the stencil. */
struct imp_sparse_reg;
int imp_toep_set_sparse_reg(imp_toep *op, const struct imp_sparse_reg *reg);
/* Pass-E schedule: coils summed per workgroup (more coils per group = fewer
 * partial images for pass F, fewer workgroups).  imp_toep_create picks one
 * that fills the GPU; set accepts [that value, ncoils] (the autotuner's
 * knob, impatient/batch.py). */
int imp_toep_get_schedule(imp_toep *op, int *coils_per_group);
int imp_toep_set_schedule(imp_toep *op, int coils_per_group);
/* Replace the field map (device [N], rad/s, NULL = 0); Q_l and the
 * preconditioner do not depend on it, so one plan serves every field map of
 * a trajectory (fMRI frames, training draws).  No counterpart in the
 * reference, which rebuilds everything per call (main_mri.cu:359-620). */
int imp_toep_set_fieldmap(imp_toep *op, const float *fm);
/* out = A p (device pointers, [N] each). */
int imp_toep_apply(imp_toep *op, const float *p_r, const float *p_i,
                   float *out_r, float *out_i);

typedef struct {
    int    iterations;        /* CG iterations performed                     */
    int    converged;         /* recursive residual reached tol              */
    int    best_iteration;    /* iteration of the returned best iterate      */
    double rel_residual;      /* sqrt(beta2/|b|^2) at exit (recursive)       */
    double best_rel_residual; /* same, for the returned iterate              */
    double true_rel_residual; /* ||A x - b|| / ||b|| recomputed (flag)       */
    double seconds;           /* wall time of the solve                      */
    int    start_used;        /* 1: started from x0; 0: from zero            */
    double start_rel_residual;/* ||b - A x0|| / ||b|| of the given x0 (-1: none) */
} imp_cg_stats;

#define IMP_CG_BEST_ITERATE   0x1  /* return argmin-residual iterate (recon.cu:2107-2110) */
#define IMP_CG_USE_X0         0x2  /* start from x instead of 0                 */
#define IMP_CG_TRUE_RESIDUAL  0x4  /* recompute ||Ax-b||/||b|| after the solve  */
#define IMP_CG_PRECOND        0x8  /* preconditioned CG with the operator's
                                      circulant M^-1 (imp_toep_set_precond)   */
#define IMP_CG_FALLBACK_ZERO  0x10 /* with USE_X0: start from 0 instead when
                                      ||b - A x0|| >= ||b|| (stats report it)  */
#define IMP_CG_GRAPH          0x20 /* capture one iteration (every kernel of the
                                      apply, dots, scalar updates, M^-1) into a
                                      HIP graph and replay it per iteration    */

/* CG on A x = b (toeplitz_recon's loop, recon.cu:2011-2133).  All scalars
 * (dot products, alpha, beta) are FP64 and stay on the device; tol is on
 * the relative recursive residual sqrt(|r|^2/|b|^2). Device pointers. */
int imp_toep_cg(imp_toep *op, const float *b_r, const float *b_i,
                float *x_r, float *x_i, int max_iter, double tol, int flags,
                imp_cg_stats *stats);

/* Circulant preconditioner of a Toeplitz operator: M^-1 = F^H diag(w) F
 * with F the DFT of the N1 x N2 x N3 image grid (periodic), w[k] > 0 in
 * natural DFT order, including the 1/(N1 N2 N3) of the inverse DFT, so M is
 * Hermitian positive definite and PCG stays valid.  imp_toep_create sets w
 * to T. Chan's optimal circulant approximation of sum_l Q_l plus lambda
 * times the periodic Laplacian spectrum.  set: copy w (device, N1*N2*N3
 * floats, every entry > 0 and finite; NULL restores the Chan weights);
 * get: copy the current weights out (device).  No counterpart in the
 * reference, whose CG is unpreconditioned (recon.cu:2011-2133). */
int imp_toep_set_precond(imp_toep *op, const float *w);
int imp_toep_get_precond(imp_toep *op, float *w);
/* w = w_chan * g(rho): Chan weights times a radial gain profile g given at
 * nbins equally spaced radii rho = |k / N| in [0, sqrt(dims)/2] (host array,
 * piecewise constant per bin, every gain > 0 and finite). */
int imp_toep_set_precond_radial(imp_toep *op, const float *gains, int nbins);
/* z = M^-1 r (device, interleaved Re/Im split as r_r/r_i -> z_r/z_i). */
int imp_toep_precond_apply(imp_toep *op, const float *r_r, const float *r_i, float *z_r,
                           float *z_i);

/* ------------------------------------------------ learned warm start
 * A 6-layer 3x3 residual CNN (32 channels, bf16 weights and activations,
 * FP32 accumulation on v_mfma_f32_32x32x16_bf16) with its own inference and
 * training kernels.  Input z is a linear estimate of the solution (e.g.
 * M^-1 b); with s = rms(z) the network returns x = z + s * net(z / s).
 * Images are H x W (device, split Re/Im, row-major), W a multiple of 64.
 * Parameters: imp_cnn_num_params() floats, per layer [9][32 co][32 ci]
 * weights then 32 biases (FP32 masters; bf16 copies feed the MFMA).
 * No counterpart in the reference (its CG starts from zero, recon.cu:1871). */
typedef struct imp_cnn imp_cnn;
int imp_cnn_create(imp_ctx *ctx, int H, int W, unsigned long long seed, imp_cnn **out);
int imp_cnn_destroy(imp_cnn *net);
int imp_cnn_num_params(void);
int imp_cnn_get_params(imp_cnn *net, float *host, int n);
int imp_cnn_set_params(imp_cnn *net, const float *host, int n);
/* x = z + s * net(z / s) */
int imp_cnn_forward(imp_cnn *net, const float *z_r, const float *z_i, float *x_r, float *x_i);
/* loss = 0.5 |x - t|^2 / |t|^2 (x as in imp_cnn_forward) and its gradient
 * with respect to the parameters (host, imp_cnn_num_params floats, or NULL) */
int imp_cnn_gradient(imp_cnn *net, const float *z_r, const float *z_i, const float *t_r,
                     const float *t_i, float *grad_host, double *loss);
/* one Adam step (beta 0.9 / 0.999) on the pair (z, t) */
int imp_cnn_train_step(imp_cnn *net, const float *z_r, const float *z_i, const float *t_r,
                       const float *t_i, float lr, double *loss);

/* ------------------------------------------ sparse-matrix regularizer
 * R = C^T W C for a real sparse C (the -reg option and its Matrix Market
 * file c.rmtx).  The reference converts C and C^T to CSR and applies
 * Cf = C p, Cf[block b] *= W, lambda C^T Cf each iteration
 * (toeplitz/recon.cu:1543-1640, bruteForce/bruteForceGpu.cu:309-360,
 * 513-541, 579-609); here both CSR copies stay resident on the device.
 * W (device [num_cols], NULL = identity) multiplies row r of C by
 * W[r mod num_cols], i.e. every block of num_cols rows. */
typedef struct imp_sparse_reg imp_sparse_reg;
/* Matrix Market "coordinate real general" reader (readMtxFile,
 * kernels/mmio.cpp:668-830): sizes first, then up to cap entries, 0-based.
 * Host only; no GPU is touched. */
int imp_read_mtx_size(const char *path, int *num_rows, int *num_cols, int *num_nonzeros);
int imp_read_mtx(const char *path, int cap, int *I, int *J, float *V);
/* C from host COO arrays (0-based; duplicates add; every index is checked). */
int imp_sparse_reg_create(imp_ctx *ctx, int num_rows, int num_cols, int num_nonzeros,
                          const int *I, const int *J, const float *V, imp_sparse_reg **out);
int imp_sparse_reg_destroy(imp_sparse_reg *reg);
/* out = lambda C^T W C p (device pointers, [num_cols] each). */
int imp_sparse_reg_apply(imp_sparse_reg *reg, float lambda, const float *w,
                         const float *p_r, const float *p_i, float *out_r, float *out_i);

/* ---------------------------------------------- brute-force CG (a11)
 * bruteForceGpu semantics (bruteForce/bruteForceGpu.cu:142-813): Fessler CG
 * on (F^H F + lambda R) x = F^H y starting from x, cg_num - 1 iterations,
 * alpha = Re(d^H g) / (|F d|^2 + Re((lambda R d)^H d)). Device pointers;
 * x_r/x_i [num_x] in/out. lambda <= 0 disables the regularizer.
 * Image is Nimg x Nimg (square, as the reference assumes). */
int imp_bruteforce_cg(imp_ctx *ctx, int num_k, int num_x, int ncoils,
                      const float *kx, const float *ky, const float *t,
                      const float *x, const float *y, const float *fm,
                      const float *d_r, const float *d_i,
                      const float *s_r, const float *s_i,
                      int cg_num, float lambda,
                      float *x_r, float *x_i);
/* The same with the brute force's other options (bruteForceGpu.cu:143-156):
 * reg != NULL: R = C^T W C from a sparse-matrix regularizer created on ctx
 * (the -reg branch, bruteForceGpu.cu:513-541, 579-629; C is (2 num_x) x num_x);
 * tv_num > 1: total-variation passes (bruteForceGpu.cu:431-757), each
 * restarting the CG from the current image with W = 0.5 / sqrt(|D_h x|^2 +
 * |D_v x|^2 + 1e-6) of it (W = 1 in the first pass); tv_num <= 1: one pass. */
int imp_bruteforce_cg_ex(imp_ctx *ctx, int num_k, int num_x, int ncoils,
                         const float *kx, const float *ky, const float *t,
                         const float *x, const float *y, const float *fm,
                         const float *d_r, const float *d_i,
                         const float *s_r, const float *s_i,
                         int cg_num, float lambda, const imp_sparse_reg *reg,
                         int tv_num, float *x_r, float *x_i);

/* ------------------------------------------------------------ data files
 * Reference file formats (kernels/structures.cpp:718-1110,
 * toeplitz/utils.cpp:520-905): v0.2 = float version, int per_coil, int
 * ncoils, int nslices, int total, float[total]; v1.0 = text "key = value"
 * header, "//" comments, per-coil "Binary_Size = n" + "Binary:" sections of
 * raw float32, then "EOF".  Host-only; no GPU needed. */
typedef struct {
    float version;      /* 0.2 or 1.0                                  */
    int   per_coil;     /* elements per coil                           */
    int   ncoils;       /* coil_number (1 if absent)                   */
    int   nslices;
    int   total;        /* file_size / total                           */
    int   nx, ny, nz;   /* v1.0 xDimension/yDimension/zDimension (-1)  */
} imp_file_info;

/* Read header only (data == NULL) or header + data into caller buffer
 * of capacity `cap` floats. */
int imp_read_data_file(const char *path, imp_file_info *info,
                       float *data, int cap);
int imp_write_data_file_v02(const char *path, int per_coil, int ncoils,
                            int nslices, const float *data, int total);
/* ncoils = -1 writes no coil_number key, as writeDataFile_JGAI_10 does for
 * coil-independent files; otherwise one Binary section per coil. */
int imp_write_data_file_v10(const char *path, int nx, int ny, int nz,
                            int ncoils, int nslices, const float *data,
                            int total);

/* ------------------------------------ reference-named file-level wrappers
 * Same arguments and meaning as the reference functions; host pointers;
 * the process-wide context on device `imp_set_default_device()` (0).
 * They return IMP_* codes instead of exiting. */
int imp_set_default_device(int device);

/* toeplitz_computeQ_GPU (toeplitz/computeQ.cmem.cuh:55-61). Reads
 * kx,ky,kz,t,ixQ,iyQ,izQ,phiR,phiI from dir; Qr[l], Qi[l] caller-allocated
 * (l = 0..ntime_segs), each of the Q-grid size.  Direct evaluation only
 * (enable_gridding is IMP_EUNSUPPORTED).  Writes Q_file when writeQ. */
int imp_toeplitz_computeQ_GPU(const char *dir, float ntime_segs, int Nx,
                              int Ny, int Nz, const char *Q_file,
                              float **Qr, float **Qi, int enable_direct,
                              int enable_gridding, float gridOS,
                              int enable_writeQ);

/* toeplitz_computeFH_GPU (toeplitz/computeFH.cmem.cuh:55-60). */
int imp_toeplitz_computeFH_GPU(const char *dir, float ntime_segs, int Nx,
                               int Ny, int Nz, const char *FhD_file,
                               float *outR, float *outI, int enable_direct,
                               int enable_gridding, float gridOS);

/* Sparse matrix in coordinate form, 0-based, HOST arrays: the reference's
 * CooMatrix (kernels/structures.h:927-960) that loadInputData fills from
 * c.rmtx (structures.cpp:598, 722-737). */
typedef struct {
    int num_rows, num_cols, num_nonzeros;
    const int *I;      /* row indices    [num_nonzeros] */
    const int *J;      /* column indices [num_nonzeros] */
    const float *V;    /* values         [num_nonzeros] */
} imp_coo_matrix;

/* toeplitz_recon (toeplitz/recon.cuh:55): argv[1..17] = N1 N2 N3
 * numRestarts numIterMax symTraj lambda2 D Dp R W Qfile FhDfile dir tv_num
 * ntime_segs outfile (main_mri.cu:578-598).  enable_reg: R = C^T W C with
 * C from coo, or from <dir>/c.rmtx when coo is NULL (C is (ndim N) x N);
 * otherwise the finite-difference stencil. */
int imp_toeplitz_recon(int argc, char **argv, int enable_tv, int enable_reg,
                       int enable_fd, int enable_reuseQ, const imp_coo_matrix *coo,
                       float *Fhd_R, float *Fhd_I, float **Q_R, float **Q_I);

/* bruteForceGpu for one slice (bruteForce/gpuPrototypes.cuh:90-105, def
 * bruteForceGpu.cu:143), host pointers: x_r/x_i [num_i] in/out (initial
 * estimate in, image out); y_r/y_i [ncoils*num_k]; kx,ky,t [num_k];
 * ix,iy,fm [num_i]; s_r/s_i [ncoils*num_i] or NULL (S = 1).  fd_penalizer
 * > 0 enables the finite-difference term (the -fd/-fdp options).  No
 * power-of-two padding. */
int imp_bruteForceGpu(float *x_r, float *x_i, const float *y_r, const float *y_i,
                      const float *kx, const float *ky, const float *ix,
                      const float *iy, const float *fm, const float *t,
                      int cg_num, int num_k, int num_i, float fd_penalizer,
                      const float *s_r, const float *s_i, int ncoils);
/* The same with bruteForceGpu's other options (bruteForceGpu.cu:143-156):
 * c != NULL: the sparse-matrix regularizer fd_penalizer C^T W C in place of
 * the stencil (enable_regularization with CooMatrix *c; c is
 * (2 num_i) x num_i, host arrays); enable_tv: tv_num total-variation passes
 * (enable_total_variation, tv_num; see imp_bruteforce_cg_ex). */
int imp_bruteForceGpu_ex(float *x_r, float *x_i, const float *y_r, const float *y_i,
                         const float *kx, const float *ky, const float *ix,
                         const float *iy, const float *fm, const float *t,
                         int cg_num, int num_k, int num_i, float fd_penalizer,
                         const float *s_r, const float *s_i, int ncoils,
                         const imp_coo_matrix *c, int enable_tv, int tv_num);
/* The same with bruteForceGpu's enable_tv_update (bruteForceGpu.cu:143-156,
 * 452-463): with enable_tv, hook (if not NULL) is called on the calling
 * thread at the start of every TV pass, tv_pass = 0 .. tv_num - 1, with the
 * image that pass starts from (host arrays of num_x floats, valid during the
 * call).  The reference writes that image to its output files there; the
 * mriSolver -tv_update option does the same from the hook. */
typedef void (*imp_tv_pass_hook)(void *user, int tv_pass, const float *x_r, const float *x_i,
                                 int num_x);
int imp_bruteForceGpu_tv(float *x_r, float *x_i, const float *y_r, const float *y_i,
                         const float *kx, const float *ky, const float *ix,
                         const float *iy, const float *fm, const float *t,
                         int cg_num, int num_k, int num_i, float fd_penalizer,
                         const float *s_r, const float *s_i, int ncoils,
                         const imp_coo_matrix *c, int enable_tv, int tv_num,
                         imp_tv_pass_hook hook, void *hook_user);

#ifdef __cplusplus
}
#endif
#endif /* IMPATIENT_H */
