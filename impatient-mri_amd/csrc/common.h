// Internal helpers shared by the HIP translation units of libimpatient.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <set>
#include <string>
#include <vector>

#include "impatient.h"

namespace imp {

// ---------------------------------------------------------------- errors
// Thread-local last error; every C entry point records its failure here.
void set_error(const std::string &msg);
void clear_error();
const char *last_error();

struct Status {
    int code = IMP_OK;
    Status() = default;
    Status(int c) : code(c) {}
    bool ok() const { return code == IMP_OK; }
};

int fail(int code, const char *fmt, ...);

#define IMP_HIP(call)                                                        \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess)                                                \
            return ::imp::fail(IMP_EHIP, "%s:%d %s: %s", __FILE__, __LINE__, \
                               #call, hipGetErrorString(e_));                \
    } while (0)

#define IMP_TRY(expr)                 \
    do {                              \
        int s_ = (expr);              \
        if (s_ != IMP_OK) return s_;  \
    } while (0)

#define IMP_CHECK_LAUNCH() IMP_HIP(hipGetLastError())

// --------------------------------------------------------------- context
struct Workspace {
    void *ptr = nullptr;
    size_t bytes = 0;
};

// One timed launch: HIP events recorded on the context stream around it.
struct ProfRec {
    const char *name;
    hipEvent_t begin, end;
};

// Independent scratch regions of a context (a call may use several at once).
enum WsSlot {
    WS_SPOS = 0,   // NUDFT stream positions
    WS_LANES,      // NUDFT lane positions
    WS_PARTIAL,    // NUDFT chunk partial sums
    WS_BPACK,      // NUDFT packed weights
    WS_W,          // NUDFT complex weights
    WS_MAXBITS,    // NUDFT weight scale
    WS_BFORCE,     // brute-force CG vectors
    WS_REG,        // regularizer / TV staging
    WS_QHALF,      // Q on the half grid (Hermitian fill)
    WS_BATCH,      // batched NUDFT operands
    WS_QS_SORT,    // separable Q: sample order, bucket counts, weight scale
    WS_QS_TAB,     // separable Q: phasor and weight tables
    WS_QS_PART,    // separable Q: K-block partial sums and job table
    WS_STACK,      // stacked-trajectory detection flags
    WS_QSTACK,     // stacked Q: one partition's 2D planes and the z factor
    WS_GRID,       // gridding: binned samples, grids and line-transform tables
    WS_COUNT
};

}  // namespace imp

struct imp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 256;
    // growable scratch regions; index selects an independent region
    imp::Workspace ws[imp::WS_COUNT];
    // FP64 reduction scratch for dot products (device)
    double *dscratch = nullptr;
    size_t dscratch_n = 0;
    // pinned host words for scalar readback
    double *host_scalars = nullptr;
    // FFT twiddle tables: table for length L at offset L (L = 2..2048)
    float2 *twiddle = nullptr;
    // live Toeplitz operators created on this context
    std::set<struct imp_toep *> toeps;
    // live networks created on this context
    std::set<struct imp_cnn *> cnns;
    // live sparse-matrix regularizers created on this context
    std::set<struct imp_sparse_reg *> sregs;
    // kernel timer (imp_profile_*): off by default
    bool prof_on = false;
    std::vector<imp::ProfRec> prof_recs;   // recorded this session
    std::vector<hipEvent_t> prof_free;     // reusable events
};

namespace imp {

// Returns a device scratch region of at least `bytes` (grows; shrinks only
// through workspace_release).
int workspace(imp_ctx *ctx, int slot, size_t bytes, void **out);
// Frees a slot's region if it holds more than `keep` bytes (synchronises the
// context stream first).
int workspace_release(imp_ctx *ctx, int slot, size_t keep);
int ensure_dscratch(imp_ctx *ctx, size_t n);

// Frees a network's device memory and detaches it from its context (for
// imp_ctx_destroy; a later imp_cnn_destroy then frees only the host object).
void detach_cnn(struct imp_cnn *net);
// The same for a sparse-matrix regularizer (imp_sparse_reg_destroy then
// frees only the host object).
void detach_sparse_reg(struct imp_sparse_reg *reg);

// Kernel timer: prof_begin records an event before a launch and returns a
// record index (-1 when the timer is off); prof_end records the closing event.
int prof_begin(imp_ctx *ctx, const char *name);
void prof_end(imp_ctx *ctx, int rec);

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// ---------------------------------------------------------- device maths
// sin and cos of 2*pi*r for r in revolutions.  The AMD transcendental unit
// takes its argument in revolutions (v_sin_f32(x) = sin(2 pi x)), so the
// only range reduction needed is r - rint(r), done exactly by the caller.
__device__ __forceinline__ void sincos_rev(float r, float *s, float *c) {
    *s = __builtin_amdgcn_sinf(r);
    *c = __builtin_amdgcn_cosf(r);
}

// r - nearest integer; exact in FP32 for |r| < 2^22.
__device__ __forceinline__ float frac_near(float r) { return r - __builtin_rintf(r); }

// Two-sum: s + e == a + b exactly.
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e) {
    s = a + b;
    float bp = s - a;
    e = (a - (s - bp)) + (b - bp);
}

struct cfloat { float x, y; };

}  // namespace imp
