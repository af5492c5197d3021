// mriSolver: command-line driver with the reference's options and file
// behaviour (main.cpp:103-452 option parsing, main_mri.cu:207-630 Toeplitz
// driver, main_mri.cu:652-1025 brute-force driver), on top of the C ABI of
// libimpatient.so.
//
//   mriSolver -idir D [-odir O] [-gpu_id n] [-cg_num K]
//             [-toeplitzDirect | -toeplitzGridding] [-ntime_segs L]
//             [-writeQ] [-reuseQ dir] [-fd] [-fdp lambda] [-tv] [-tv_num n]
//             [-reg] [-cpu] [-nogpu] [-mgpu] [-gridOS_Q f] [-gridOS_FH f]
//             [-version] [-help]
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "impatient.h"

namespace {

struct Options {
    std::string idir, odir;
    bool have_idir = false, have_odir = false;
    int cg_num = 8;                 // CG_NUM_DEFAULT (main.cpp:96)
    int tv_num = 10;                // TV_NUM_DEFAULT
    float fdp = 0.0f;               // FDP_DEFAULT
    bool fdp_given = false;
    bool toeplitz_direct = false, toeplitz_gridding = false;
    float ntime_segs = 12.0f;       // NTIME_SEG_DEFAULT (main.cpp:101)
    float gridOS_Q = 1.125f, gridOS_FHD = 1.5f;
    int gpu_id = 0;
    bool reuseQ = false, writeQ = false;
    std::string reuse_loc;
    bool fd = false, tv = false, tv_update = false, reg = false;
    bool cpu = false, nogpu = false, mgpu = false;
};

void usage(FILE *f) {
    std::fprintf(f,
                 "Usage: mriSolver -idir <input dir> [options]\n"
                 "  -odir <dir>          output directory (default <idir>/output)\n"
                 "  -gpu_id <n>          GPU device (default 0)\n"
                 "  -cg_num <n>          CG iterations (default 8)\n"
                 "  -toeplitzDirect      Toeplitz method, direct Q and F^H d\n"
                 "  -toeplitzGridding    Toeplitz method, Q and F^H d by gridding\n"
                 "  -gridOS_Q f / -gridOS_FH f  gridding oversampling in [1, 2] (default 1.125 / 1.5)\n"
                 "  -ntime_segs <L>      time segments (default 12), after the method flag\n"
                 "  -writeQ              write <idir>/Q_stone.file\n"
                 "  -reuseQ <dir>        read <dir>/Q_stone.file instead of computing Q\n"
                 "  -fd                  finite-difference regularizer (needs -fdp)\n"
                 "  -fdp <lambda>        regularization weight (enables -fd)\n"
                 "  -tv / -tv_num <n>    total-variation outer iterations (-tv_num enables -tv)\n"
                 "  -tv_update           export the brute-force image before every TV pass\n"
                 "  -reg                 sparse-matrix regularizer C^T W C, C from <idir>/c.rmtx\n"
                 "  -cpu / -nogpu / -mgpu  CPU and multi-GPU paths (not part of this engine)\n"
                 "  -version, -help\n");
}

bool need_arg(int i, int argc, const char *flag) {
    if (i + 1 >= argc) {
        std::fprintf(stderr, "Error: option %s needs a value\n", flag);
        return false;
    }
    return true;
}

int parse(int argc, char **argv, Options &o) {
    for (int i = 1; i < argc; ++i) {
        const char *a = argv[i];
        if (!std::strcmp(a, "-idir")) {
            if (!need_arg(i, argc, a)) return 1;
            o.idir = argv[++i];
            o.have_idir = true;
        } else if (!std::strcmp(a, "-odir")) {
            if (!need_arg(i, argc, a)) return 1;
            o.odir = argv[++i];
            o.have_odir = true;
        } else if (!std::strcmp(a, "-gpu_id")) {
            if (!need_arg(i, argc, a)) return 1;
            o.gpu_id = std::atoi(argv[++i]);
        } else if (!std::strcmp(a, "-cg_num")) {
            if (!need_arg(i, argc, a)) return 1;
            o.cg_num = std::atoi(argv[++i]);
        } else if (!std::strcmp(a, "-toeplitzDirect")) {
            o.toeplitz_direct = true;
        } else if (!std::strcmp(a, "-toeplitzGridding")) {
            o.toeplitz_gridding = true;
        } else if (!std::strcmp(a, "-ntime_segs")) {
            // must follow the strategy flag (main.cpp:226-240)
            if (!o.toeplitz_direct && !o.toeplitz_gridding) {
                std::fprintf(stderr, "Error: -ntime_segs must follow -toeplitzDirect or "
                                     "-toeplitzGridding\n");
                return 1;
            }
            if (!need_arg(i, argc, a)) return 1;
            o.ntime_segs = (float)std::atof(argv[++i]);
            if (o.ntime_segs < 1.0f) {
                std::fprintf(stderr, "\nError: number of time segments has to be a natural "
                                     "number in floating format.\n");
                return 1;
            }
        } else if (!std::strcmp(a, "-gridOS_Q") || !std::strcmp(a, "-gridOS_FH")) {
            // gridding oversampling of Q / of F^H d, after -toeplitzGridding,
            // in [1, 2] (main.cpp:256-290)
            if (!o.toeplitz_gridding || !need_arg(i, argc, a) || argv[i + 1][0] == '-') {
                std::fprintf(stderr, "\nError: -toeplitzGridding flag is not found or invalid "
                                     "gridding OS factor.\n");
                return 1;
            }
            const float v = (float)std::atof(argv[++i]);
            if (v < 1.0f || v > 2.0f) {
                std::fprintf(stderr, "\nError: gridding oversampling factor should be in "
                                     "[1.0,2.0].\n");
                return 1;
            }
            (a[8] == 'Q' ? o.gridOS_Q : o.gridOS_FHD) = v;
        } else if (!std::strcmp(a, "-reuseQ")) {
            if (!o.toeplitz_direct && !o.toeplitz_gridding) {
                std::fprintf(stderr, "Error: -reuseQ must follow the Toeplitz method flag\n");
                return 1;
            }
            if (!need_arg(i, argc, a)) return 1;
            o.reuseQ = true;
            o.reuse_loc = argv[++i];
        } else if (!std::strcmp(a, "-writeQ")) {
            o.writeQ = true;
        } else if (!std::strcmp(a, "-fd")) {
            o.fd = true;
        } else if (!std::strcmp(a, "-fdp")) {
            // setting the weight enables the finite-difference term (main.cpp:292-301)
            if (!need_arg(i, argc, a) || argv[i + 1][0] == '-') {
                std::fprintf(stderr, "\nError: Missing finite difference penalty.\n");
                return 1;
            }
            o.fdp = (float)std::atof(argv[++i]);
            o.fdp_given = true;
            o.fd = true;
        } else if (!std::strcmp(a, "-tv")) {
            o.tv = true;
        } else if (!std::strcmp(a, "-tv_num")) {
            // setting the pass count enables TV (main.cpp:306-319)
            if (!need_arg(i, argc, a) || argv[i + 1][0] == '-') {
                std::fprintf(stderr, "\nError: Missing TV iteration number.\n");
                return 1;
            }
            o.tv_num = std::atoi(argv[++i]);
            if (o.tv_num == 0) {
                std::fprintf(stderr, "\nError: Zero or wrong TV iteration number format.\n");
                return 1;
            }
            o.tv = true;
        } else if (!std::strcmp(a, "-tv_update")) {
            o.tv = true;
            o.tv_update = true;
        } else if (!std::strcmp(a, "-reg")) {
            o.reg = true;
        } else if (!std::strcmp(a, "-cpu")) {
            o.cpu = true;
        } else if (!std::strcmp(a, "-nogpu")) {
            o.nogpu = true;
        } else if (!std::strcmp(a, "-mgpu")) {
            o.mgpu = true;
        } else if (!std::strcmp(a, "-version")) {
            std::printf("%s\n", imp_version_string());
            std::exit(0);
        } else if (!std::strcmp(a, "-help")) {
            usage(stdout);
            std::exit(0);
        } else {
            std::fprintf(stderr, "\nError: Unknown option \"%s\"\n", a);
            usage(stderr);
            return 1;
        }
    }
    if (argc == 1 || !o.have_idir) {
        std::fprintf(stderr, "\nError: Missing input directory name.\n");
        usage(stderr);
        return 1;
    }
    if ((o.fd || o.tv) && !o.fdp_given) {
        std::fprintf(stderr, "\n Error: Regularization enabled, but no weight parameter "
                             "specified. Hint: -fdp [lambda].\n");
        return 1;
    }
    if (!o.have_odir) o.odir = o.idir + "/output";
    return 0;
}

int report(int st, const char *what) {
    if (st != IMP_OK)
        std::fprintf(stderr, "Error: %s: %s (%s)\n", what, imp_status_string(st), imp_last_error());
    return st;
}

struct Arr {
    std::vector<float> v;
    imp_file_info info{};
};

int load(const std::string &path, Arr &a) {
    int st = imp_read_data_file(path.c_str(), &a.info, nullptr, 0);
    if (st != IMP_OK) return report(st, path.c_str());
    a.v.assign(a.info.total > 0 ? a.info.total : 1, 0.f);
    return report(imp_read_data_file(path.c_str(), &a.info, a.v.data(), (int)a.v.size()),
                  path.c_str());
}

int write_like(const std::string &path, float version, int nx, int ny, int nz, int per_coil,
               int ncoils, const std::vector<float> &data) {
    if (version < 1.0f)
        return report(imp_write_data_file_v02(path.c_str(), per_coil, ncoils, 1, data.data(),
                                              (int)data.size()),
                      path.c_str());
    return report(imp_write_data_file_v10(path.c_str(), nx, ny, nz, ncoils, 1, data.data(),
                                          (int)data.size()),
                  path.c_str());
}

// The Toeplitz driver of main_mri.cu:207-630.
int run_toeplitz(const Options &o) {
    // direct evaluation when -toeplitzDirect is given, else gridding (main_mri.cu:207-630)
    const int direct = o.toeplitz_direct ? 1 : 0, grid = direct ? 0 : 1;
    Arr kx, ix, iy, kd;
    if (load(o.idir + "/kx.dat", kx) || load(o.idir + "/ix.dat", ix) ||
        load(o.idir + "/iy.dat", iy) || load(o.idir + "/kdata_r.dat", kd))
        return IMP_EIO;
    const float version = kx.info.version;
    const int num_k = kx.info.per_coil, num_i = ix.info.per_coil;
    const int ncoils = kd.info.ncoils > 0 ? kd.info.ncoils : 1;
    int Nz = (version >= 1.0f && ix.info.nz > 0) ? ix.info.nz : 1;
    const int Ny = (int)std::lround(std::sqrt((double)num_i / Nz));
    const int Nx = Ny;
    if ((long)Nx * Ny * Nz != num_i) {
        std::fprintf(stderr, "Error: image size %d is not Ny*Nx*Nz with Nx = Ny\n", num_i);
        return IMP_EINVAL;
    }
    std::printf("  Image data size (yxz)  : (%d,%d,%d)\n  K-space data size/coil : %d\n"
                "  Number of coils        : %d\n  Number of time segments: %f\n",
                Ny, Nx, Nz, num_k, ncoils, o.ntime_segs);
    // phiR = 1, phiI = 0 written into the input directory (main_mri.cu:364-385)
    {
        const int n = num_k * ncoils;
        std::vector<float> one(n, 1.0f), zero(n, 0.0f);
        if (write_like(o.idir + "/phiR.dat", version, Nx, Ny, Nz, num_k, ncoils, one) ||
            write_like(o.idir + "/phiI.dat", version, Nx, Ny, Nz, num_k, ncoils, zero))
            return IMP_EIO;
    }
    const int L = (int)o.ntime_segs;
    const int M3 = Nz > 1 ? 2 * Nz : 1;
    const int nq = 4 * Nx * Ny * M3;
    std::vector<std::vector<float>> Qr(L + 1, std::vector<float>(nq)),
        Qi(L + 1, std::vector<float>(nq));
    std::vector<float *> pr(L + 1), pi(L + 1);
    for (int l = 0; l <= L; ++l) {
        pr[l] = Qr[l].data();
        pi[l] = Qi[l].data();
    }
    if (o.reuseQ) {
        // size check of main_mri.cu:400-450
        const std::string qf = o.reuse_loc + "/Q_stone.file";
        struct stat sb;
        if (lstat(qf.c_str(), &sb) < 0) {
            std::fprintf(stderr, "Error: bad directory path that contains the Q file\n");
            return IMP_EIO;
        }
        const long long want = (long long)(L + 1) * nq * 4 * 2;
        if ((long long)sb.st_size != want) {
            std::fprintf(stderr, "Error: incorrect Q file size\n");
            return IMP_EFORMAT;
        }
    } else {
        // Q-grid coordinates, ordered like the image (H2): y outer / x inner
        // when the image varies fastest along x (main_mri.cu:461-500), x outer
        // / y inner when the coordinate files vary fastest along y.
        const bool y_fast = num_i > 1 && ix.v[1] == ix.v[0] && iy.v[1] != iy.v[0];
        std::vector<float> gx(nq), gy(nq), gz(nq, 0.f);
        int idx = 0;
        for (int a = -Ny; a < Ny; ++a)
            for (int b = -Nx; b < Nx; ++b)
                for (int c = (Nz > 1 ? -Nz : 0); c < (Nz > 1 ? Nz : 1); ++c) {
                    const float yv = (float)(y_fast ? b : a) / (float)Ny;
                    const float xv = (float)(y_fast ? a : b) / (float)Nx;
                    gy[idx] = yv;
                    gx[idx] = xv;
                    if (Nz > 1) gz[idx] = (float)c / (float)Nz;
                    ++idx;
                }
        const int Nxq = 2 * Nx, Nyq = 2 * Ny, Nzq = M3;
        auto wq = [&](const char *name, const std::vector<float> &v) {
            const std::string p = o.idir + "/" + name;
            if (version < 1.0f)
                return report(imp_write_data_file_v02(p.c_str(), nq, ncoils, 1, v.data(), nq),
                              p.c_str());
            return report(imp_write_data_file_v10(p.c_str(), Nxq, Nyq, Nzq, -1, 1, v.data(), nq),
                          p.c_str());
        };
        if (wq("ixQ.dat", gx) || wq("iyQ.dat", gy) || wq("izQ.dat", gz)) return IMP_EIO;
        std::printf("======== Step 1. Compute Q matrices: Q[%d] ========\n", L + 1);
        const std::string qfile = o.idir + "/Q_stone.file";
        if (report(imp_toeplitz_computeQ_GPU(o.idir.c_str(), o.ntime_segs, Nx, Ny, Nz,
                                             qfile.c_str(), pr.data(), pi.data(), direct, grid,
                                             o.gridOS_Q, o.writeQ ? 1 : 0),
                   "computeQ"))
            return IMP_EHIP;
    }
    std::printf("======== Step 2. Compute F^H(d)  ========\n");
    std::vector<float> fr(num_i), fi(num_i);
    const std::string fhd = o.idir + "/FhD.file";
    if (report(imp_toeplitz_computeFH_GPU(o.idir.c_str(), o.ntime_segs, Nx, Ny, Nz, fhd.c_str(),
                                          fr.data(), fi.data(), direct, grid, o.gridOS_FHD),
               "computeFH"))
        return IMP_EHIP;
    std::printf("======== Step 3. Start CG solver ========\n");
    char buf[18][4096];
    const std::string qloc = (o.reuseQ ? o.reuse_loc : o.idir) + "/Q_stone.file";
    std::snprintf(buf[0], sizeof buf[0], "toeplitz_recon");
    std::snprintf(buf[1], sizeof buf[1], "%d", Ny);
    std::snprintf(buf[2], sizeof buf[2], "%d", Nx);
    std::snprintf(buf[3], sizeof buf[3], "%d", Nz);
    std::snprintf(buf[4], sizeof buf[4], "%d", 1);
    std::snprintf(buf[5], sizeof buf[5], "%d", o.cg_num);
    std::snprintf(buf[6], sizeof buf[6], "%d", 0);
    std::snprintf(buf[7], sizeof buf[7], "%f", o.fdp);
    std::snprintf(buf[8], sizeof buf[8], "D.file");
    std::snprintf(buf[9], sizeof buf[9], "Dp.file");
    std::snprintf(buf[10], sizeof buf[10], "R.file");
    std::snprintf(buf[11], sizeof buf[11], "W.file");
    std::snprintf(buf[12], sizeof buf[12], "%s", qloc.c_str());
    std::snprintf(buf[13], sizeof buf[13], "%s/FhD.file", o.idir.c_str());
    std::snprintf(buf[14], sizeof buf[14], "%s", o.idir.c_str());
    std::snprintf(buf[15], sizeof buf[15], "%d", o.tv_num);
    std::snprintf(buf[16], sizeof buf[16], "%f", o.ntime_segs);
    std::snprintf(buf[17], sizeof buf[17], "%s/out.file", o.idir.c_str());
    char *argv[18];
    for (int i = 0; i < 18; ++i) argv[i] = buf[i];
    if (report(imp_toeplitz_recon(18, argv, o.tv ? 1 : 0, o.reg ? 1 : 0, o.fd ? 1 : 0,
                                  o.reuseQ ? 1 : 0, nullptr, fr.data(), fi.data(), pr.data(),
                                  pi.data()),
               "recon"))
        return IMP_EHIP;
    std::printf("  GPU: Exporting results to %s\n", buf[17]);
    return IMP_OK;
}

// output_gpu_{r,i}.dat: the first n values of xr / xi as text "%f\n"
// (exportDataCpu, structures.cpp:1166-1175)
int export_text(const std::string &fr, const std::string &fi, const std::vector<float> &xr,
                const std::vector<float> &xi, size_t n) {
    FILE *f1 = std::fopen(fr.c_str(), "wt"), *f2 = std::fopen(fi.c_str(), "wt");
    if (!f1 || !f2) {
        std::fprintf(stderr, "Error: cannot write %s\n", (f1 ? fi : fr).c_str());
        if (f1) std::fclose(f1);
        if (f2) std::fclose(f2);
        return IMP_EIO;
    }
    for (size_t i = 0; i < n; ++i) {
        std::fprintf(f1, "%f\n", xr[i]);
        std::fprintf(f2, "%f\n", xi[i]);
    }
    std::fclose(f1);
    std::fclose(f2);
    return IMP_OK;
}

// The brute-force driver of main_mri.cu:652-1025 (GPU part).
int run_bruteforce(const Options &o) {
    const char *names[] = {"kx", "ky", "ix", "iy", "t", "fm", "kdata_r", "kdata_i",
                           "idata_r", "idata_i", "sensi_r", "sensi_i"};
    Arr a[12];
    for (int i = 0; i < 12; ++i)
        if (load(o.idir + "/" + names[i] + ".dat", a[i])) return IMP_EIO;
    const int num_k = a[0].info.per_coil, num_i = a[2].info.per_coil;
    const int ncoils = a[6].info.ncoils > 0 ? a[6].info.ncoils : 1;
    // slices: one brute-force solve each, one after another on this GPU
    // (main_mri.cu:889-905); slice s starts at s*num_i in idata and fm and at
    // s*num_k in kdata (the reference's offset, which does not scale with
    // the coil count)
    const int nslices = a[6].info.nslices > 1 ? a[6].info.nslices : 1;
    const size_t need_x = (size_t)nslices * num_i, need_k = (size_t)(nslices - 1) * num_k +
                                                            (size_t)ncoils * num_k;
    if (a[8].v.size() < need_x || a[9].v.size() < need_x || a[5].v.size() < need_x ||
        a[6].v.size() < need_k || a[7].v.size() < need_k) {
        std::fprintf(stderr, "Error: %d slices need %zu image and %zu k-space values\n", nslices,
                     need_x, need_k);
        return IMP_EFORMAT;
    }
    std::vector<float> xr(a[8].v.begin(), a[8].v.begin() + need_x),
        xi(a[9].v.begin(), a[9].v.begin() + need_x);
    const float lambda = (o.fd || o.tv || o.reg) ? o.fdp : 0.0f;
    int st = IMP_OK;
    imp_coo_matrix c{};
    std::vector<int> ri, rj;
    std::vector<float> rv;
    if (o.reg) {
        // C from <idir>/c.rmtx (loadInputData, structures.cpp:598, 722-737)
        const std::string fn = o.idir + "/c.rmtx";
        st = imp_read_mtx_size(fn.c_str(), &c.num_rows, &c.num_cols, &c.num_nonzeros);
        if (report(st, "read c.rmtx")) return st;
        ri.resize((size_t)c.num_nonzeros + 1);
        rj.resize((size_t)c.num_nonzeros + 1);
        rv.resize((size_t)c.num_nonzeros + 1);
        st = imp_read_mtx(fn.c_str(), c.num_nonzeros, ri.data(), rj.data(), rv.data());
        if (report(st, "read c.rmtx")) return st;
        c.I = ri.data();
        c.J = rj.data();
        c.V = rv.data();
    }
    mkdir(o.odir.c_str(), 0755);
    const std::string fr = o.odir + "/output_gpu_r.dat", fi = o.odir + "/output_gpu_i.dat";
    // -tv_update: before every TV pass the image is exported (bruteForceGpu.cu:452-463);
    // the files hold every slice, this slice's current image in its place
    struct Update {
        std::vector<float> *xr, *xi;
        size_t offset, need_x;
        const std::string *fr, *fi;
        int st;
    } upd{&xr, &xi, 0, need_x, &fr, &fi, IMP_OK};
    const imp_tv_pass_hook hook = [](void *user, int, const float *x_r, const float *x_i,
                                     int n) {
        Update *u = (Update *)user;
        std::vector<float> r(*u->xr), im(*u->xi);
        std::copy(x_r, x_r + n, r.begin() + u->offset);
        std::copy(x_i, x_i + n, im.begin() + u->offset);
        if (export_text(*u->fr, *u->fi, r, im, u->need_x) != IMP_OK) u->st = IMP_EIO;
    };
    for (int sl = 0; sl < nslices && st == IMP_OK; ++sl) {
        if (nslices > 1) std::printf("  GPU: Processing slice %d.\n", sl);
        float *x_r = xr.data() + (size_t)sl * num_i, *x_i = xi.data() + (size_t)sl * num_i;
        const float *y_r = a[6].v.data() + (size_t)sl * num_k,
                    *y_i = a[7].v.data() + (size_t)sl * num_k;
        const float *fm = a[5].v.data() + (size_t)sl * num_i;
        const float *s_r = ncoils > 1 ? a[10].v.data() : nullptr,
                    *s_i = ncoils > 1 ? a[11].v.data() : nullptr;
        upd.offset = (size_t)sl * num_i;
        // TV passes as bruteForceGpu runs them (bruteForceGpu.cu:431-757)
        st = imp_bruteForceGpu_tv(x_r, x_i, y_r, y_i, a[0].v.data(), a[1].v.data(),
                                  a[2].v.data(), a[3].v.data(), fm, a[4].v.data(), o.cg_num,
                                  num_k, num_i, lambda, s_r, s_i, ncoils, o.reg ? &c : nullptr,
                                  o.tv ? 1 : 0, o.tv_num, o.tv_update ? hook : nullptr, &upd);
        if (st == IMP_OK && upd.st != IMP_OK) st = upd.st;
    }
    if (report(st, "bruteForceGpu")) return st;
    // output_gpu_{r,i}.dat, text "%f\n" (exportDataCpu, structures.cpp:1166-1175):
    // every slice, num_slices * num_i values (main_mri.cu:917-918)
    if (export_text(fr, fi, xr, xi, need_x) != IMP_OK) return IMP_EIO;
    std::printf("  GPU: Exporting results to %s\n", o.odir.c_str());
    return IMP_OK;
}

}  // namespace

int main(int argc, char **argv) {
    Options o;
    if (parse(argc, argv, o)) return 1;
    if (o.cpu || o.nogpu)
        std::fprintf(stderr, "mriSolver: the CPU path is not part of this engine; running on "
                             "the GPU only.\n");
    if (o.mgpu) std::fprintf(stderr, "mriSolver: -mgpu ignored (one GPU).\n");
    imp_set_default_device(o.gpu_id);
    const int st = (o.toeplitz_direct || o.toeplitz_gridding) ? run_toeplitz(o) : run_bruteforce(o);
    if (st != IMP_OK) {
        std::fprintf(stderr, "Error: Failed to execute the mriSolver.\n");
        return 1;
    }
    return 0;
}
