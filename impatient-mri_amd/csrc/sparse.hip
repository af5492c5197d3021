// Sparse-matrix regularizer lambda C^T W C (the -reg option).
//
// The reference reads a real Matrix Market matrix C (c.rmtx, N columns,
// ndim x N rows: one block of differences per image axis), converts it and
// its transpose to CSR on the host, and for every CG iteration runs
//   Cf = C p;  Cf[block b] *= W (b < ndim);  out += lambda C^T Cf
// as two scalar CSR products per real/imaginary part, a pointwise product
// per block and an add (toeplitz/recon.cu:1543-1640; the brute-force CG does
// the same at bruteForce/bruteForceGpu.cu:513-541, 579-609).
//
// Here both CSR copies stay resident for the life of the object, the
// complex vector is read interleaved (one 8-byte load per nonzero for both
// parts), W is folded into the first product and lambda and the
// accumulation into the second: two launches per application, one thread
// per row.  Regularizer matrices have a handful of nonzeros per row, so a
// row per lane keeps the CSR arrays streaming and the gathers of p mostly
// within a few cache lines of each other (difference stencils).
#include "common.h"
#include "toeplitz.h"

#include <vector>

namespace imp {

namespace {

// cf[r] = W[r mod n] * sum_e val[e] p[idx[e]],  e in [ptr[r], ptr[r+1]).
// blockIdx.y is the block of n rows, so W's index needs no division.
__global__ void k_csr_wc(int rows, int n, const int *__restrict__ ptr,
                         const int *__restrict__ idx, const float *__restrict__ val,
                         const float2 *__restrict__ p, const float *__restrict__ w,
                         float2 *__restrict__ cf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const long rl = (long)blockIdx.y * n + i;
    if (i >= n || rl >= rows) return;
    const int r = (int)rl;
    float sx = 0.f, sy = 0.f;
    const int e1 = ptr[r + 1];
    for (int e = ptr[r]; e < e1; ++e) {
        const float v = val[e];
        const float2 q = p[idx[e]];
        sx = fmaf(v, q.x, sx);
        sy = fmaf(v, q.y, sy);
    }
    if (w) {
        const float ww = w[i];
        sx *= ww;
        sy *= ww;
    }
    cf[r] = make_float2(sx, sy);
}

// out[i] (+)= lambda * sum_e val[e] cf[idx[e]]  (rows of C^T)
__global__ void k_csr_t(int rows, const int *__restrict__ ptr, const int *__restrict__ idx,
                        const float *__restrict__ val, const float2 *__restrict__ cf,
                        float lambda, float2 *__restrict__ out, int accumulate) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    float sx = 0.f, sy = 0.f;
    const int e1 = ptr[i + 1];
    for (int e = ptr[i]; e < e1; ++e) {
        const float v = val[e];
        const float2 q = cf[idx[e]];
        sx = fmaf(v, q.x, sx);
        sy = fmaf(v, q.y, sy);
    }
    float2 o = make_float2(lambda * sx, lambda * sy);
    if (accumulate) {
        const float2 a = out[i];
        o.x += a.x;
        o.y += a.y;
    }
    out[i] = o;
}

// CSR of the nnz entries (row[e], col[e], v[e]) with `rows` rows: a stable
// counting sort by row, so entries of one row keep their file order
// (duplicates stay separate terms of the row sum, as in coo2Csr,
// structures.cpp:380-392, which does not compact them here).
void build_csr(int rows, long nnz, const int *row, const int *col, const float *v,
               std::vector<int> &ptr, std::vector<int> &idx, std::vector<float> &val) {
    ptr.assign((size_t)rows + 1, 0);
    for (long e = 0; e < nnz; ++e) ++ptr[(size_t)row[e] + 1];
    for (int r = 0; r < rows; ++r) ptr[(size_t)r + 1] += ptr[r];
    std::vector<int> next(ptr.begin(), ptr.end() - 1);
    idx.resize((size_t)nnz);
    val.resize((size_t)nnz);
    for (long e = 0; e < nnz; ++e) {
        const int at = next[row[e]]++;
        idx[at] = col[e];
        val[at] = v[e];
    }
}

}  // namespace

int sparse_reg(imp_ctx *ctx, const SparseReg &R, float lambda, const float2 *p, const float *w,
               float2 *out, bool accumulate) {
    int rec = prof_begin(ctx, "k_csr_wc");
    hipLaunchKernelGGL(k_csr_wc, dim3(ceil_div(R.cols, 256), ceil_div(R.rows, R.cols)), dim3(256),
                       0, ctx->stream, R.rows, R.cols, R.c_ptr, R.c_idx, R.c_val, p, w, R.cf);
    IMP_CHECK_LAUNCH();
    prof_end(ctx, rec);
    rec = prof_begin(ctx, "k_csr_t");
    hipLaunchKernelGGL(k_csr_t, dim3(ceil_div(R.cols, 256)), dim3(256), 0, ctx->stream, R.cols,
                       R.t_ptr, R.t_idx, R.t_val, R.cf, lambda, out, accumulate ? 1 : 0);
    IMP_CHECK_LAUNCH();
    prof_end(ctx, rec);
    return IMP_OK;
}

}  // namespace imp

using namespace imp;

struct imp_sparse_reg {
    imp_ctx *ctx = nullptr;
    SparseReg R;
    std::vector<void *> owned;
};

namespace imp {
void detach_sparse_reg(imp_sparse_reg *reg) {
    for (void *p : reg->owned) (void)hipFree(p);
    reg->owned.clear();
    reg->R = SparseReg{};
    reg->ctx = nullptr;
}
}  // namespace imp

// After imp_ctx_destroy the object is detached (device memory already freed,
// no context): only the host object is freed here.  Operators that still use
// it lose it, and their applies fail until imp_toep_set_sparse_reg.
extern "C" int imp_sparse_reg_destroy(imp_sparse_reg *reg) {
    if (!reg) return IMP_OK;
    if (reg->ctx) {
        (void)hipStreamSynchronize(reg->ctx->stream);
        for (imp_toep *op : reg->ctx->toeps)
            if (op->P.sreg == &reg->R) {
                op->P.sreg = nullptr;
                op->P.sreg_lost = true;
            }
        reg->ctx->sregs.erase(reg);
        detach_sparse_reg(reg);
    }
    delete reg;
    return IMP_OK;
}

extern "C" int imp_sparse_reg_create(imp_ctx *ctx, int num_rows, int num_cols, int num_nonzeros,
                                     const int *I, const int *J, const float *V,
                                     imp_sparse_reg **out) {
    if (out) *out = nullptr;
    if (!ctx || !out || num_rows <= 0 || num_cols <= 0 || num_nonzeros < 0 ||
        (num_nonzeros > 0 && (!I || !J || !V)))
        return fail(IMP_EINVAL, "imp_sparse_reg_create: bad arguments");
    // one launch row of the first product per block of num_cols rows
    if (ceil_div(num_rows, num_cols) > 65535)
        return fail(IMP_EINVAL, "imp_sparse_reg_create: %d x %d has more than 65535 blocks of %d "
                                "rows", num_rows, num_cols, num_cols);
    // every index is checked here, so the kernels' gathers stay in bounds
    for (long e = 0; e < num_nonzeros; ++e) {
        if (I[e] < 0 || I[e] >= num_rows || J[e] < 0 || J[e] >= num_cols)
            return fail(IMP_EINVAL,
                        "imp_sparse_reg_create: entry %ld (%d, %d) outside the %d x %d matrix", e,
                        I[e], J[e], num_rows, num_cols);
    }
    std::vector<int> cp, ci, tp, ti;
    std::vector<float> cv, tv;
    build_csr(num_rows, num_nonzeros, I, J, V, cp, ci, cv);
    build_csr(num_cols, num_nonzeros, J, I, V, tp, ti, tv);
    imp_sparse_reg *reg = new imp_sparse_reg;
    reg->ctx = ctx;
    ctx->sregs.insert(reg);
    SparseReg &R = reg->R;
    R.rows = num_rows;
    R.cols = num_cols;
    R.nnz = num_nonzeros;
    auto up = [&](const void *src, size_t bytes, void **dst) -> int {
        if (hipMalloc(dst, bytes ? bytes : 4) != hipSuccess) {
            (void)hipGetLastError();
            return fail(IMP_ENOMEM, "imp_sparse_reg_create: cannot allocate %zu bytes", bytes);
        }
        reg->owned.push_back(*dst);
        if (src && bytes) IMP_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        return IMP_OK;
    };
    int st = IMP_OK;
    const size_t nz = (size_t)num_nonzeros;
    if (st == IMP_OK) st = up(cp.data(), cp.size() * sizeof(int), (void **)&R.c_ptr);
    if (st == IMP_OK) st = up(ci.data(), nz * sizeof(int), (void **)&R.c_idx);
    if (st == IMP_OK) st = up(cv.data(), nz * sizeof(float), (void **)&R.c_val);
    if (st == IMP_OK) st = up(tp.data(), tp.size() * sizeof(int), (void **)&R.t_ptr);
    if (st == IMP_OK) st = up(ti.data(), nz * sizeof(int), (void **)&R.t_idx);
    if (st == IMP_OK) st = up(tv.data(), nz * sizeof(float), (void **)&R.t_val);
    // the work vector W C p, one entry per row of C
    if (st == IMP_OK) st = up(nullptr, (size_t)num_rows * sizeof(float2), (void **)&R.cf);
    if (st != IMP_OK) {
        imp_sparse_reg_destroy(reg);
        return st;
    }
    *out = reg;
    return IMP_OK;
}

extern "C" int imp_sparse_reg_apply(imp_sparse_reg *reg, float lambda, const float *w,
                                    const float *p_r, const float *p_i, float *out_r,
                                    float *out_i) {
    if (!reg || !p_r || !p_i || !out_r || !out_i)
        return fail(IMP_EINVAL, "imp_sparse_reg_apply: bad arguments");
    if (!reg->ctx) return fail(IMP_EINVAL, "imp_sparse_reg_apply: regularizer has no context");
    imp_ctx *ctx = reg->ctx;
    const long n = reg->R.cols;
    void *buf = nullptr;
    IMP_TRY(workspace(ctx, WS_REG, 2 * n * sizeof(float2), &buf));
    float2 *pc = (float2 *)buf, *sc = pc + n;
    IMP_TRY(split_to_c(ctx, p_r, p_i, pc, n));
    IMP_TRY(sparse_reg(ctx, reg->R, lambda, pc, w, sc, false));
    IMP_TRY(c_to_split(ctx, sc, out_r, out_i, n));
    return IMP_OK;
}

namespace imp {
// nullptr for NULL and for a regularizer whose context was destroyed
const SparseReg *sparse_reg_of(const imp_sparse_reg *reg) {
    return reg && reg->ctx ? &reg->R : nullptr;
}
imp_ctx *sparse_reg_ctx(const imp_sparse_reg *reg) { return reg ? reg->ctx : nullptr; }
}  // namespace imp
