// Reference-named, file-level entry points (host pointers, synchronous):
// drop-in replacements for the functions main_mri.cu calls
// (toeplitz_computeQ_GPU, toeplitz_computeFH_GPU, toeplitz_recon,
// bruteForceGpu).  They read the same files from the same directory and
// return IMP_* codes where the reference calls exit(1).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "toeplitz.h"

using namespace imp;

namespace {

std::mutex g_mu;
imp_ctx *g_ctx = nullptr;
int g_device = 0;
int g_compat = 0;

int default_ctx(imp_ctx **out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ctx) IMP_TRY(imp_ctx_create(g_device, &g_ctx));
    IMP_HIP(hipSetDevice(g_ctx->device));
    *out = g_ctx;
    return IMP_OK;
}

struct HostArray {
    std::vector<float> v;
    imp_file_info info{};
};

int read_file(const std::string &path, HostArray &a) {
    IMP_TRY(imp_read_data_file(path.c_str(), &a.info, nullptr, 0));
    a.v.resize(a.info.total > 0 ? a.info.total : 1);
    IMP_TRY(imp_read_data_file(path.c_str(), &a.info, a.v.data(), (int)a.v.size()));
    return IMP_OK;
}

bool file_exists(const std::string &p) {
    FILE *f = std::fopen(p.c_str(), "rb");
    if (!f) return false;
    std::fclose(f);
    return true;
}

struct Dev {
    float *p = nullptr;
    ~Dev() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t n) {
        IMP_HIP(hipMalloc(&p, (n ? n : 1) * sizeof(float)));
        return IMP_OK;
    }
    int up(const float *h, size_t n) {
        IMP_TRY(alloc(n));
        IMP_HIP(hipMemcpy(p, h, n * sizeof(float), hipMemcpyHostToDevice));
        return IMP_OK;
    }
    int up(const std::vector<float> &h, size_t n) { return up(h.data(), n); }
};

int write_raw(const char *path, const std::vector<const float *> &parts, size_t n) {
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(IMP_EIO, "cannot open %s for writing", path);
    for (const float *p : parts)
        if (std::fwrite(p, sizeof(float), n, f) != n) {
            std::fclose(f);
            return fail(IMP_EIO, "%s: write failed", path);
        }
    std::fclose(f);
    return IMP_OK;
}

// tau and t0 exactly as recon.cu:1785-1787 / computeFH.cmem.cu:1226-1229
void segment_timing(const std::vector<float> &t, int nk, float ntime_segs, float &tau, float &t0) {
    tau = ntime_segs > 0.0f ? (t[nk - 1] - t[0]) / ntime_segs : (t[nk - 1] - t[0]);
    t0 = t[0];
}

// 2D arrays in the reference's y-fastest file order ([x][y], the MATLAB
// meshgrid order of mriData/64x64x1) <-> the engine's [y][x] (x fastest).
// `to_engine` converts file order to engine order; planes of n0 x n1 with
// n0 = Nx, n1 = Ny.
void transpose_plane(std::vector<float> &v, size_t off, int Nx, int Ny, bool to_engine) {
    std::vector<float> tmp(v.begin() + off, v.begin() + off + (size_t)Nx * Ny);
    for (int y = 0; y < Ny; ++y)
        for (int x = 0; x < Nx; ++x) {
            const size_t e = (size_t)y * Nx + x, f = (size_t)x * Ny + y;
            v[off + (to_engine ? e : f)] = tmp[to_engine ? f : e];
        }
}

// the samples must lie in the window [-N/2, N/2] on every axis gridded
// (computeQ.cmem.cu:527-534)
int check_window(const HostArray &kx, const HostArray &ky, const HostArray &kz, int nk, int Nx,
                 int Ny, int Nz) {
    for (int i = 0; i < nk; ++i)
        if (std::fabs(kx.v[i]) > Nx / 2.0f || std::fabs(ky.v[i]) > Ny / 2.0f ||
            (Nz > 1 && std::fabs(kz.v[i]) > Nz / 2.0f))
            return fail(IMP_EINVAL, "gridding: k-space sample %d outside the window [-N/2, N/2]",
                        i);
    return IMP_OK;
}

// the -toeplitzGridding branch of toeplitz_computeQ_GPU (computeQ.cmem.cu:401-754)
int computeQ_gridding(const char *dir, float ntime_segs, int Nx, int Ny, int Nz,
                      const char *Q_file, float **Qr, float **Qi, float gridOS, int writeQ) {
    const std::string d(dir);
    HostArray kx, ky, kz, t, gx, gy;
    IMP_TRY(read_file(d + "/kx.dat", kx));
    IMP_TRY(read_file(d + "/ky.dat", ky));
    IMP_TRY(read_file(d + "/kz.dat", kz));
    IMP_TRY(read_file(d + "/t.dat", t));
    const int nk = kx.info.per_coil;
    if (nk <= 0 || Nx <= 0 || Ny <= 0 || Nz <= 0 || (int)ky.v.size() < nk ||
        (int)kz.v.size() < nk || (int)t.v.size() < nk)
        return fail(IMP_EFORMAT, "imp_toeplitz_computeQ_GPU: inconsistent file sizes in %s", dir);
    IMP_TRY(check_window(kx, ky, kz, nk, Nx, Ny, Nz));
    // the Q-grid order of the caller's coordinate files (H2): 2D grids whose
    // x is constant over the first entries hold y fastest
    bool y_fast = false;
    if (Nz == 1 && file_exists(d + "/ixQ.dat") && file_exists(d + "/iyQ.dat")) {
        IMP_TRY(read_file(d + "/ixQ.dat", gx));
        IMP_TRY(read_file(d + "/iyQ.dat", gy));
        y_fast = gx.v.size() > 1 && gx.v[1] == gx.v[0] && gy.v[1] != gy.v[0];
    }
    imp_ctx *ctx = nullptr;
    IMP_TRY(default_ctx(&ctx));
    const int L = (int)ntime_segs;
    const long nseg = L + 1;
    const long nq = 4L * Nx * Ny * (Nz > 1 ? 2L * Nz : 1L);
    Dev dkx, dky, dkz, dt, dq;
    IMP_TRY(dkx.up(kx.v, nk));
    IMP_TRY(dky.up(ky.v, nk));
    IMP_TRY(dkz.up(kz.v, nk));
    IMP_TRY(dt.up(t.v, nk));
    IMP_TRY(dq.alloc(2 * nseg * nq));
    IMP_TRY(gridding_q(ctx, nk, dkx.p, dky.p, Nz > 1 ? dkz.p : nullptr, dt.p, Nx, Ny, Nz, L,
                       gridOS, dq.p, dq.p + nseg * nq));
    IMP_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<float> h(2 * nseg * nq);
    IMP_HIP(hipMemcpy(h.data(), dq.p, h.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (long l = 0; l < nseg; ++l) {
        if (!Qr[l] || !Qi[l]) return fail(IMP_EINVAL, "imp_toeplitz_computeQ_GPU: Qr[%ld] NULL", l);
        if (y_fast) {
            transpose_plane(h, l * nq, 2 * Nx, 2 * Ny, false);
            transpose_plane(h, (nseg + l) * nq, 2 * Nx, 2 * Ny, false);
        }
        std::memcpy(Qr[l], h.data() + l * nq, nq * sizeof(float));
        std::memcpy(Qi[l], h.data() + (nseg + l) * nq, nq * sizeof(float));
    }
    if (writeQ && Q_file) {
        // (L+1) x [Re(Q_l), Im(Q_l)] float32 (computeQ.cmem.cu:280-300)
        FILE *f = std::fopen(Q_file, "wb");
        if (!f) return fail(IMP_EIO, "cannot open %s for writing", Q_file);
        for (long l = 0; l < nseg; ++l)
            if (std::fwrite(Qr[l], sizeof(float), nq, f) != (size_t)nq ||
                std::fwrite(Qi[l], sizeof(float), nq, f) != (size_t)nq) {
                std::fclose(f);
                return fail(IMP_EIO, "%s: write failed", Q_file);
            }
        std::fclose(f);
    }
    return IMP_OK;
}

// the -toeplitzGridding branch of toeplitz_computeFH_GPU
// (parcomputeAH_GPU_Grid, computeFH.cmem.cu:1032-1117, 1330-1336)
int computeFH_gridding(const char *dir, float ntime_segs, int Nx, int Ny, int Nz,
                       const char *FhD_file, float *outR, float *outI, float gridOS) {
    const std::string d(dir);
    HostArray kx, ky, kz, t, ix, iy, fm, sr, si, dr, di;
    IMP_TRY(read_file(d + "/kx.dat", kx));
    IMP_TRY(read_file(d + "/ky.dat", ky));
    IMP_TRY(read_file(d + "/kz.dat", kz));
    IMP_TRY(read_file(d + "/t.dat", t));
    IMP_TRY(read_file(d + "/ix.dat", ix));
    IMP_TRY(read_file(d + "/iy.dat", iy));
    IMP_TRY(read_file(d + "/fm.dat", fm));
    IMP_TRY(read_file(d + "/kdata_r.dat", dr));
    IMP_TRY(read_file(d + "/kdata_i.dat", di));
    const int nk = kx.info.per_coil;
    const long nv = (long)Nx * Ny * Nz;
    const int nc = dr.info.ncoils > 0 ? dr.info.ncoils : 1;
    if (nk <= 0 || nv <= 0 || (long)ix.info.per_coil != nv || (long)fm.v.size() < nv ||
        (long)dr.v.size() < (long)nk * nc || (long)di.v.size() < (long)nk * nc ||
        (int)t.v.size() < nk || (int)kz.v.size() < nk)
        return fail(IMP_EFORMAT, "imp_toeplitz_computeFH_GPU: inconsistent file sizes in %s", dir);
    IMP_TRY(check_window(kx, ky, kz, nk, Nx, Ny, Nz));
    const bool sense = nc > 1;   // single-coil data: S = 1 (computeFH.cmem.cu:305-310)
    if (sense) {
        IMP_TRY(read_file(d + "/sensi_r.dat", sr));
        IMP_TRY(read_file(d + "/sensi_i.dat", si));
        if ((long)sr.v.size() < nv * nc || (long)si.v.size() < nv * nc)
            return fail(IMP_EFORMAT, "imp_toeplitz_computeFH_GPU: sensitivity files too short");
    }
    // image order of the coordinate files (H2): the gridded image is [y][x]
    const bool y_fast = Nz == 1 && nv > 1 && ix.v[1] == ix.v[0] && iy.v[1] != iy.v[0];
    if (y_fast) {
        transpose_plane(fm.v, 0, Nx, Ny, true);
        for (int c = 0; sense && c < nc; ++c) {
            transpose_plane(sr.v, (size_t)c * nv, Nx, Ny, true);
            transpose_plane(si.v, (size_t)c * nv, Nx, Ny, true);
        }
    }
    imp_ctx *ctx = nullptr;
    IMP_TRY(default_ctx(&ctx));
    Dev a[10];
    IMP_TRY(a[0].up(kx.v, nk));
    IMP_TRY(a[1].up(ky.v, nk));
    IMP_TRY(a[2].up(kz.v, nk));
    IMP_TRY(a[3].up(t.v, nk));
    IMP_TRY(a[4].up(fm.v, nv));
    IMP_TRY(a[5].up(dr.v, (size_t)nk * nc));
    IMP_TRY(a[6].up(di.v, (size_t)nk * nc));
    if (sense) {
        IMP_TRY(a[7].up(sr.v, (size_t)nv * nc));
        IMP_TRY(a[8].up(si.v, (size_t)nv * nc));
    }
    IMP_TRY(a[9].alloc(2 * (size_t)nv));
    IMP_TRY(gridding_fhd(ctx, nk, nc, a[0].p, a[1].p, Nz > 1 ? a[2].p : nullptr, a[3].p, Nx, Ny,
                         Nz, a[4].p, a[5].p, a[6].p, a[7].p, a[8].p, (int)ntime_segs, gridOS,
                         a[9].p, a[9].p + nv));
    IMP_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<float> h(2 * nv);
    IMP_HIP(hipMemcpy(h.data(), a[9].p, h.size() * sizeof(float), hipMemcpyDeviceToHost));
    if (y_fast) {
        transpose_plane(h, 0, Nx, Ny, false);
        transpose_plane(h, nv, Nx, Ny, false);
    }
    std::memcpy(outR, h.data(), nv * sizeof(float));
    std::memcpy(outI, h.data() + nv, nv * sizeof(float));
    // FhD.file: Re then Im float32 (computeFH.cmem.cu:313-321)
    if (FhD_file) IMP_TRY(write_raw(FhD_file, {outR, outI}, nv));
    return IMP_OK;
}

}  // namespace

extern "C" int imp_set_compat(int flags) {
    g_compat = flags;
    return IMP_OK;
}
extern "C" int imp_get_compat(void) { return g_compat; }

extern "C" int imp_set_default_device(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ctx && g_ctx->device != device) {
        imp_ctx_destroy(g_ctx);
        g_ctx = nullptr;
    }
    g_device = device;
    return IMP_OK;
}

extern "C" int imp_toeplitz_computeQ_GPU(const char *dir, float ntime_segs, int Nx, int Ny,
                                         int Nz, const char *Q_file, float **Qr, float **Qi,
                                         int enable_direct, int enable_gridding, float gridOS,
                                         int enable_writeQ) {
    if (!dir || !Qr || !Qi || ntime_segs < 0)
        return fail(IMP_EINVAL, "imp_toeplitz_computeQ_GPU: bad arguments");
    if (!enable_direct && !enable_gridding)
        return fail(IMP_EINVAL, "imp_toeplitz_computeQ_GPU: neither direct nor gridding");
    if (!enable_direct) return computeQ_gridding(dir, ntime_segs, Nx, Ny, Nz, Q_file, Qr, Qi,
                                                 gridOS, enable_writeQ);
    const std::string d(dir);
    HostArray kx, ky, kz, t, gx, gy, gz, pr, pi;
    IMP_TRY(read_file(d + "/kx.dat", kx));
    IMP_TRY(read_file(d + "/ky.dat", ky));
    IMP_TRY(read_file(d + "/kz.dat", kz));
    IMP_TRY(read_file(d + "/t.dat", t));
    IMP_TRY(read_file(d + "/ixQ.dat", gx));
    IMP_TRY(read_file(d + "/iyQ.dat", gy));
    IMP_TRY(read_file(d + "/izQ.dat", gz));
    const bool have_phi = file_exists(d + "/phiR.dat") && file_exists(d + "/phiI.dat");
    if (have_phi) {
        IMP_TRY(read_file(d + "/phiR.dat", pr));
        IMP_TRY(read_file(d + "/phiI.dat", pi));
    }
    const int nk = kx.info.per_coil;
    const int nq = gx.info.per_coil;
    if (nk <= 0 || nq <= 0 || (int)t.v.size() < nk || (int)ky.v.size() < nk ||
        (int)gy.v.size() < nq || (int)gz.v.size() < nq)
        return fail(IMP_EFORMAT, "imp_toeplitz_computeQ_GPU: inconsistent file sizes in %s", dir);
    const long expect = 4L * Nx * Ny * (Nz > 1 ? 2L * Nz : 1L);
    if (Nx > 0 && Ny > 0 && nq != expect)
        return fail(IMP_EFORMAT, "imp_toeplitz_computeQ_GPU: Q grid has %d points, expected %ld",
                    nq, expect);
    imp_ctx *ctx = nullptr;
    IMP_TRY(default_ctx(&ctx));
    const int L = (int)ntime_segs;
    const long nseg = L + 1;
    Dev dkx, dky, dkz, dt, dgx, dgy, dgz, dpr, dpi, dq;
    IMP_TRY(dkx.up(kx.v, nk));
    IMP_TRY(dky.up(ky.v, nk));
    IMP_TRY(dkz.up(kz.v, nk));
    IMP_TRY(dt.up(t.v, nk));
    IMP_TRY(dgx.up(gx.v, nq));
    IMP_TRY(dgy.up(gy.v, nq));
    IMP_TRY(dgz.up(gz.v, nq));
    if (have_phi) {
        IMP_TRY(dpr.up(pr.v, nk));
        IMP_TRY(dpi.up(pi.v, nk));
    }
    IMP_TRY(dq.alloc(2 * nseg * nq));
    const float tt[2] = {t.v[0], t.v[nk - 1]};
    const int flags = (g_compat & IMP_COMPAT_Q_CUMULATIVE) ? IMP_Q_CUMULATIVE : 0;
    IMP_TRY(toeplitz_q(ctx, nk, dkx.p, dky.p, Nz > 1 ? dkz.p : nullptr, dt.p,
                       have_phi ? dpr.p : nullptr, have_phi ? dpi.p : nullptr, Nx, Ny, Nz, nq,
                       dgx.p, dgy.p, Nz > 1 ? dgz.p : nullptr, L, flags, tt, dq.p,
                       dq.p + nseg * nq));
    IMP_HIP(hipStreamSynchronize(ctx->stream));
    for (int l = 0; l <= L; ++l) {
        if (!Qr[l] || !Qi[l]) return fail(IMP_EINVAL, "imp_toeplitz_computeQ_GPU: Qr[%d] NULL", l);
        IMP_HIP(hipMemcpy(Qr[l], dq.p + (long)l * nq, nq * sizeof(float), hipMemcpyDeviceToHost));
        IMP_HIP(hipMemcpy(Qi[l], dq.p + (nseg + l) * (long)nq, nq * sizeof(float),
                          hipMemcpyDeviceToHost));
    }
    if (enable_writeQ && Q_file) {
        // (L+1) x [Re(Q_l), Im(Q_l)] float32 (computeQ.cmem.cu:280-300)
        FILE *f = std::fopen(Q_file, "wb");
        if (!f) return fail(IMP_EIO, "cannot open %s for writing", Q_file);
        for (int l = 0; l <= L; ++l)
            if (std::fwrite(Qr[l], sizeof(float), nq, f) != (size_t)nq ||
                std::fwrite(Qi[l], sizeof(float), nq, f) != (size_t)nq) {
                std::fclose(f);
                return fail(IMP_EIO, "%s: write failed", Q_file);
            }
        std::fclose(f);
    }
    return IMP_OK;
}

extern "C" int imp_toeplitz_computeFH_GPU(const char *dir, float ntime_segs, int Nx, int Ny,
                                          int Nz, const char *FhD_file, float *outR,
                                          float *outI, int enable_direct, int enable_gridding,
                                          float gridOS) {
    if (!dir || !outR || !outI)
        return fail(IMP_EINVAL, "imp_toeplitz_computeFH_GPU: bad arguments");
    if (!enable_direct && !enable_gridding)
        return fail(IMP_EINVAL, "imp_toeplitz_computeFH_GPU: neither direct nor gridding");
    if (!enable_direct)
        return computeFH_gridding(dir, ntime_segs, Nx, Ny, Nz, FhD_file, outR, outI, gridOS);
    const std::string d(dir);
    HostArray kx, ky, kz, t, ix, iy, iz, fm, pr, pi, sr, si, dr, di;
    IMP_TRY(read_file(d + "/kx.dat", kx));
    IMP_TRY(read_file(d + "/ky.dat", ky));
    IMP_TRY(read_file(d + "/kz.dat", kz));
    IMP_TRY(read_file(d + "/t.dat", t));
    IMP_TRY(read_file(d + "/ix.dat", ix));
    IMP_TRY(read_file(d + "/iy.dat", iy));
    IMP_TRY(read_file(d + "/iz.dat", iz));
    IMP_TRY(read_file(d + "/fm.dat", fm));
    IMP_TRY(read_file(d + "/kdata_r.dat", dr));
    IMP_TRY(read_file(d + "/kdata_i.dat", di));
    const int nk = kx.info.per_coil;
    const int nx = ix.info.per_coil;
    const int nc = dr.info.ncoils > 0 ? dr.info.ncoils : 1;
    if (nk <= 0 || nx <= 0 || (int)dr.v.size() < nk * nc || (int)di.v.size() < nk * nc ||
        (int)fm.v.size() < nx || (int)t.v.size() < nk)
        return fail(IMP_EFORMAT, "imp_toeplitz_computeFH_GPU: inconsistent file sizes in %s", dir);
    const bool have_phi = file_exists(d + "/phiR.dat") && file_exists(d + "/phiI.dat");
    if (have_phi) {
        IMP_TRY(read_file(d + "/phiR.dat", pr));
        IMP_TRY(read_file(d + "/phiI.dat", pi));
        if ((int)pr.v.size() < nk * nc || (int)pi.v.size() < nk * nc)
            return fail(IMP_EFORMAT, "imp_toeplitz_computeFH_GPU: phi files too short");
    }
    // single-coil data: S = 1 (computeFH.cmem.cu:305-310)
    const bool sense = nc > 1;
    if (sense) {
        IMP_TRY(read_file(d + "/sensi_r.dat", sr));
        IMP_TRY(read_file(d + "/sensi_i.dat", si));
        if ((int)sr.v.size() < nx * nc || (int)si.v.size() < nx * nc)
            return fail(IMP_EFORMAT, "imp_toeplitz_computeFH_GPU: sensitivity files too short");
    }
    imp_ctx *ctx = nullptr;
    IMP_TRY(default_ctx(&ctx));
    const bool d3 = Nz > 1;
    Dev a[16];
    IMP_TRY(a[0].up(kx.v, nk));
    IMP_TRY(a[1].up(ky.v, nk));
    IMP_TRY(a[2].up(kz.v, nk));
    IMP_TRY(a[3].up(t.v, nk));
    IMP_TRY(a[4].up(ix.v, nx));
    IMP_TRY(a[5].up(iy.v, nx));
    IMP_TRY(a[6].up(iz.v, nx));
    IMP_TRY(a[7].up(fm.v, nx));
    IMP_TRY(a[8].up(dr.v, (size_t)nk * nc));
    IMP_TRY(a[9].up(di.v, (size_t)nk * nc));
    if (have_phi) {
        IMP_TRY(a[10].up(pr.v, (size_t)nk * nc));
        IMP_TRY(a[11].up(pi.v, (size_t)nk * nc));
    }
    if (sense) {
        IMP_TRY(a[12].up(sr.v, (size_t)nx * nc));
        IMP_TRY(a[13].up(si.v, (size_t)nx * nc));
    }
    IMP_TRY(a[14].alloc(2 * (size_t)nx));
    IMP_TRY(nudft_adjoint(ctx, nk, nx, nc, a[0].p, a[1].p, d3 ? a[2].p : nullptr, a[3].p, a[4].p,
                          a[5].p, d3 ? a[6].p : nullptr, a[7].p, a[10].p, a[11].p, a[8].p,
                          a[9].p, a[12].p, a[13].p, a[14].p, a[14].p + nx));
    IMP_HIP(hipStreamSynchronize(ctx->stream));
    IMP_HIP(hipMemcpy(outR, a[14].p, nx * sizeof(float), hipMemcpyDeviceToHost));
    IMP_HIP(hipMemcpy(outI, a[14].p + nx, nx * sizeof(float), hipMemcpyDeviceToHost));
    // FhD.file: Re then Im float32 (computeFH.cmem.cu:313-321)
    if (FhD_file) IMP_TRY(write_raw(FhD_file, {outR, outI}, nx));
    return IMP_OK;
}

extern "C" int imp_toeplitz_recon(int argc, char **argv, int enable_tv, int enable_reg,
                                  int enable_fd, int enable_reuseQ, const imp_coo_matrix *coo,
                                  float *Fhd_R, float *Fhd_I, float **Q_R, float **Q_I) {
    (void)enable_fd;
    if (argc != 18 || !argv || !Fhd_R || !Fhd_I)
        return fail(IMP_EINVAL, "imp_toeplitz_recon: expected 18 arguments (recon.cu:1775)");
    const int N1 = std::atoi(argv[1]), N2 = std::atoi(argv[2]), N3 = std::atoi(argv[3]);
    const int numRestarts = std::atoi(argv[4]);
    const int numIterMax = std::atoi(argv[5]);
    const float lambda2 = (float)std::atof(argv[7]);
    const char *qfile = argv[12];
    const std::string dir(argv[14]);
    const int tv_num = std::atoi(argv[15]);
    const float ntime_segs = (float)std::atof(argv[16]);
    const char *outfile = argv[17];
    if (N1 <= 0 || N2 <= 0 || N3 <= 0 || numIterMax < 0 || ntime_segs < 0)
        return fail(IMP_EINVAL, "imp_toeplitz_recon: bad sizes in argv");
    const int L = (int)ntime_segs;
    const long nvox = (long)N1 * N2 * N3;
    const long G = 4L * N1 * N2 * (N3 > 1 ? 2L * N3 : 1L);
    // inputs that recon re-reads from the directory (recon.cu:1781)
    HostArray t, fm, sr, si, kd;
    IMP_TRY(read_file(dir + "/t.dat", t));
    IMP_TRY(read_file(dir + "/fm.dat", fm));
    IMP_TRY(read_file(dir + "/kdata_r.dat", kd));
    const int nc = kd.info.ncoils > 0 ? kd.info.ncoils : 1;
    const int nk = t.info.per_coil;
    if ((long)fm.v.size() < nvox || nk <= 0)
        return fail(IMP_EFORMAT, "imp_toeplitz_recon: fm.dat / t.dat sizes");
    const bool sense = nc > 1;
    if (sense) {
        IMP_TRY(read_file(dir + "/sensi_r.dat", sr));
        IMP_TRY(read_file(dir + "/sensi_i.dat", si));
        if ((long)sr.v.size() < nvox * nc || (long)si.v.size() < nvox * nc)
            return fail(IMP_EFORMAT, "imp_toeplitz_recon: sensitivity sizes");
    }
    float tau, t0;
    segment_timing(t.v, nk, ntime_segs, tau, t0);
    // Q: from memory, or from Q_stone.file when reusing (matrix_t::read_pack)
    std::vector<float> qh(2 * (size_t)(L + 1) * G);
    if (enable_reuseQ) {
        FILE *f = std::fopen(qfile, "rb");
        if (!f) return fail(IMP_EIO, "%s not found", qfile);
        std::fseek(f, 0, SEEK_END);
        const long sz = std::ftell(f);
        std::fseek(f, 0, SEEK_SET);
        if (sz != (long)(2 * (L + 1) * G * sizeof(float))) {
            std::fclose(f);
            return fail(IMP_EFORMAT, "%s: incorrect Q file size", qfile);
        }
        std::vector<float> raw(2 * (size_t)(L + 1) * G);
        const size_t got = std::fread(raw.data(), sizeof(float), raw.size(), f);
        std::fclose(f);
        if (got != raw.size()) return fail(IMP_EIO, "%s: short read", qfile);
        for (int l = 0; l <= L; ++l) {
            std::memcpy(&qh[(size_t)l * G], &raw[(size_t)2 * l * G], G * sizeof(float));
            std::memcpy(&qh[(size_t)(L + 1 + l) * G], &raw[(size_t)(2 * l + 1) * G],
                        G * sizeof(float));
        }
    } else {
        if (!Q_R || !Q_I) return fail(IMP_EINVAL, "imp_toeplitz_recon: Q arrays are NULL");
        for (int l = 0; l <= L; ++l) {
            std::memcpy(&qh[(size_t)l * G], Q_R[l], G * sizeof(float));
            std::memcpy(&qh[(size_t)(L + 1 + l) * G], Q_I[l], G * sizeof(float));
        }
    }
    imp_ctx *ctx = nullptr;
    IMP_TRY(default_ctx(&ctx));
    Dev dq, dfm, dsr, dsi, db, dx, dw;
    IMP_TRY(dq.up(qh, qh.size()));
    IMP_TRY(dfm.up(fm.v, nvox));
    if (sense) {
        IMP_TRY(dsr.up(sr.v, nvox * nc));
        IMP_TRY(dsi.up(si.v, nvox * nc));
    }
    std::vector<float> bh(2 * nvox);
    std::memcpy(bh.data(), Fhd_R, nvox * sizeof(float));
    std::memcpy(bh.data() + nvox, Fhd_I, nvox * sizeof(float));
    IMP_TRY(db.up(bh, bh.size()));
    IMP_TRY(dx.alloc(2 * nvox));
    IMP_HIP(hipMemset(dx.p, 0, 2 * nvox * sizeof(float)));
    IMP_TRY(dw.alloc(nvox));
    // sparse-matrix regularizer: C from the caller, or from <dir>/c.rmtx as
    // loadInputData reads it (structures.cpp:598, 722-737); one block of
    // nvox rows per image axis (recon.cu:1590-1631)
    imp_sparse_reg *reg = nullptr;
    struct RegGuard {
        imp_sparse_reg *&r;
        ~RegGuard() { imp_sparse_reg_destroy(r); }
    } reg_guard{reg};
    if (enable_reg) {
        imp_coo_matrix fc{};
        std::vector<int> ri, rj;
        std::vector<float> rv;
        const imp_coo_matrix *c = coo;
        if (!c) {
            const std::string fn = dir + "/c.rmtx";
            IMP_TRY(imp_read_mtx_size(fn.c_str(), &fc.num_rows, &fc.num_cols, &fc.num_nonzeros));
            ri.resize((size_t)fc.num_nonzeros + 1);
            rj.resize((size_t)fc.num_nonzeros + 1);
            rv.resize((size_t)fc.num_nonzeros + 1);
            IMP_TRY(imp_read_mtx(fn.c_str(), fc.num_nonzeros, ri.data(), rj.data(), rv.data()));
            fc.I = ri.data();
            fc.J = rj.data();
            fc.V = rv.data();
            c = &fc;
        }
        const long ndim = N3 > 1 ? 3 : 2;
        if ((long)c->num_cols != nvox || (long)c->num_rows != ndim * nvox)
            return fail(IMP_EINVAL, "imp_toeplitz_recon: C is %d x %d, expected %ld x %ld",
                        c->num_rows, c->num_cols, ndim * nvox, nvox);
        IMP_TRY(imp_sparse_reg_create(ctx, c->num_rows, c->num_cols, c->num_nonzeros, c->I, c->J,
                                      c->V, &reg));
    }
    imp_toep *op = nullptr;
    IMP_TRY(imp_toep_create(ctx, N1, N2, N3, nc, L, tau, t0, dq.p, dq.p + (size_t)(L + 1) * G,
                            dfm.p, sense ? dsr.p : nullptr, sense ? dsi.p : nullptr, lambda2,
                            &op));
    if (reg) {
        const int rs = imp_toep_set_sparse_reg(op, reg);
        if (rs != IMP_OK) {
            imp_toep_destroy(op);
            return rs;
        }
    }
    // x stays the last iterate across TV passes and restarts; reconBest is
    // the output (recon.cu:2013-2126, 2150-2162).  On restart 0 every pass
    // starts its best residual (and beta) from its own |b - A x|^2; later
    // restarts carry both from the previous pass, so reconBest changes only
    // where an iterate beats every earlier one (recon.cu:2043-2048, 2107-2110).
    // W stays the last TV weight across restarts (pass 0 does not reset it).
    std::vector<float> best(2 * nvox, 0.f);
    int st = IMP_OK;
    const int passes = enable_tv ? (tv_num > 0 ? tv_num : 1) : 1;
    for (int restart = 0; restart < (numRestarts > 0 ? numRestarts : 1) && st == IMP_OK;
         ++restart) {
        for (int tv_i = 0; tv_i < passes && st == IMP_OK; ++tv_i) {
            if (tv_i != 0) {
                st = imp_tv_weight(ctx, N1, N2, N3, dx.p, dx.p + nvox, dw.p);
                if (st == IMP_OK) st = imp_toep_set_weight(op, dw.p);
            }
            imp_cg_stats stats{};
            ToepPlan &P = op->P;
            const int carry = restart > 0 ? CG_CARRY_SCALARS : 0;
            if (st == IMP_OK) st = split_to_c(ctx, db.p, db.p + nvox, P.vb, nvox);
            if (st == IMP_OK) st = split_to_c(ctx, dx.p, dx.p + nvox, P.vx, nvox);
            if (st == IMP_OK)
                st = toep_cg(ctx, P, P.vb, P.vx, numIterMax, 1.0e-16,
                             IMP_CG_USE_X0 | IMP_CG_BEST_ITERATE | CG_KEEP_LAST | carry, &stats);
            if (st == IMP_OK) st = c_to_split(ctx, P.vx, dx.p, dx.p + nvox, nvox);
            if (st == IMP_OK && hipStreamSynchronize(ctx->stream) != hipSuccess)
                st = fail(IMP_EHIP, "imp_toeplitz_recon: CG stream");
            if (st == IMP_OK && stats.best_iteration > 0) {
                // reconBest is replaced only when a pass improved on its start
                std::vector<float2> hb(nvox);
                if (hipMemcpy(hb.data(), P.vbest, nvox * sizeof(float2), hipMemcpyDeviceToHost) !=
                    hipSuccess)
                    st = fail(IMP_EHIP, "imp_toeplitz_recon: copy of best iterate");
                for (long i = 0; i < nvox && st == IMP_OK; ++i) {
                    best[i] = hb[i].x;
                    best[nvox + i] = hb[i].y;
                }
            }
        }
    }
    imp_toep_destroy(op);
    if (st != IMP_OK) return st;
    // out.file: Re(reconBest) then Im(reconBest), float32 (recon.cu:2150-2165)
    return write_raw(outfile, {best.data(), best.data() + nvox}, nvox);
}

namespace {
// imp_bruteForceGpu with an optional sparse C (host COO)
int bforce_host(float *x_r, float *x_i, const float *y_r, const float *y_i, const float *kx,
                const float *ky, const float *ix, const float *iy, const float *fm,
                const float *t, int cg_num, int num_k, int num_i, float fd_penalizer,
                const float *s_r, const float *s_i, int ncoils, const imp_coo_matrix *c,
                int tv_num, imp_tv_pass_hook hook = nullptr, void *hook_user = nullptr) {
    if (!x_r || !x_i || !y_r || !y_i || !kx || !ky || !ix || !iy || !t || num_k <= 0 ||
        num_i <= 0 || ncoils <= 0 || (s_r == nullptr) != (s_i == nullptr))
        return fail(IMP_EINVAL, "imp_bruteForceGpu: bad arguments");
    imp_ctx *ctx = nullptr;
    IMP_TRY(default_ctx(&ctx));
    imp_sparse_reg *reg = nullptr;
    if (c)
        IMP_TRY(imp_sparse_reg_create(ctx, c->num_rows, c->num_cols, c->num_nonzeros, c->I, c->J,
                                      c->V, &reg));
    struct Guard {
        imp_sparse_reg *r;
        ~Guard() { imp_sparse_reg_destroy(r); }
    } guard{reg};
    const size_t nk = (size_t)num_k * ncoils;
    Dev a[12];
    IMP_TRY(a[0].up(kx, num_k));
    IMP_TRY(a[1].up(ky, num_k));
    IMP_TRY(a[2].up(t, num_k));
    IMP_TRY(a[3].up(ix, num_i));
    IMP_TRY(a[4].up(iy, num_i));
    if (fm) IMP_TRY(a[5].up(fm, num_i));
    IMP_TRY(a[6].up(y_r, nk));
    IMP_TRY(a[7].up(y_i, nk));
    if (s_r) {
        IMP_TRY(a[8].up(s_r, (size_t)num_i * ncoils));
        IMP_TRY(a[9].up(s_i, (size_t)num_i * ncoils));
    }
    IMP_TRY(a[10].up(x_r, num_i));
    IMP_TRY(a[11].up(x_i, num_i));
    IMP_TRY(bruteforce_cg(ctx, num_k, num_i, ncoils, a[0].p, a[1].p, a[2].p, a[3].p, a[4].p,
                          a[5].p, a[6].p, a[7].p, a[8].p, a[9].p, cg_num, fd_penalizer,
                          sparse_reg_of(reg), tv_num, a[10].p, a[11].p, hook, hook_user));
    IMP_HIP(hipMemcpy(x_r, a[10].p, num_i * sizeof(float), hipMemcpyDeviceToHost));
    IMP_HIP(hipMemcpy(x_i, a[11].p, num_i * sizeof(float), hipMemcpyDeviceToHost));
    return IMP_OK;
}
}  // namespace

extern "C" int imp_bruteForceGpu(float *x_r, float *x_i, const float *y_r, const float *y_i,
                                 const float *kx, const float *ky, const float *ix,
                                 const float *iy, const float *fm, const float *t, int cg_num,
                                 int num_k, int num_i, float fd_penalizer, const float *s_r,
                                 const float *s_i, int ncoils) {
    return bforce_host(x_r, x_i, y_r, y_i, kx, ky, ix, iy, fm, t, cg_num, num_k, num_i,
                       fd_penalizer, s_r, s_i, ncoils, nullptr, 1);
}

extern "C" int imp_bruteForceGpu_ex(float *x_r, float *x_i, const float *y_r, const float *y_i,
                                    const float *kx, const float *ky, const float *ix,
                                    const float *iy, const float *fm, const float *t,
                                    int cg_num, int num_k, int num_i, float fd_penalizer,
                                    const float *s_r, const float *s_i, int ncoils,
                                    const imp_coo_matrix *c, int enable_tv, int tv_num) {
    return bforce_host(x_r, x_i, y_r, y_i, kx, ky, ix, iy, fm, t, cg_num, num_k, num_i,
                       fd_penalizer, s_r, s_i, ncoils, c, enable_tv ? tv_num : 1);
}

extern "C" int imp_bruteForceGpu_tv(float *x_r, float *x_i, const float *y_r, const float *y_i,
                                    const float *kx, const float *ky, const float *ix,
                                    const float *iy, const float *fm, const float *t,
                                    int cg_num, int num_k, int num_i, float fd_penalizer,
                                    const float *s_r, const float *s_i, int ncoils,
                                    const imp_coo_matrix *c, int enable_tv, int tv_num,
                                    imp_tv_pass_hook hook, void *hook_user) {
    // the export runs only inside the TV loop (bruteForceGpu.cu:440-463)
    return bforce_host(x_r, x_i, y_r, y_i, kx, ky, ix, iy, fm, t, cg_num, num_k, num_i,
                       fd_penalizer, s_r, s_i, ncoils, c, enable_tv ? tv_num : 1,
                       enable_tv ? hook : nullptr, hook_user);
}
