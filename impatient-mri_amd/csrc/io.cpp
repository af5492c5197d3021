// Reference data-file formats (host only, no GPU needed).
//   v0.2: float version(0.2), int per_coil, int ncoils, int nslices, int total,
//         float data[total]                      (readDataFile, structures.cpp:780-865;
//                                                 readDataFile_JGAI, utils.cpp:520-567)
//   v1.0: text header "key = value" lines, "//" comments, then for every coil
//         "Binary_Size = n" and "Binary:" followed by n raw float32, then "EOF"
//                                                (readDataFile_10, structures.cpp:867-1110;
//                                                 writeDataFile_JGAI_10, utils.cpp:856-905)
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "impatient.h"

namespace imp {
int fail(int code, const char *fmt, ...);
}
using imp::fail;

namespace {

struct File {
    FILE *f = nullptr;
    explicit File(const char *p, const char *mode) : f(std::fopen(p, mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

// Reads one text line (without '\n'); false at end of file.
bool read_line(FILE *f, std::string &line) {
    line.clear();
    int ch;
    bool any = false;
    while ((ch = std::fgetc(f)) != EOF) {
        any = true;
        if (ch == '\n') return true;
        line.push_back((char)ch);
        if (line.size() > 4096) return true;  // malformed; caller rejects
    }
    return any;
}

std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && (s[a] == ' ' || s[a] == '\t' || s[a] == '\r')) ++a;
    while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t' || s[b - 1] == '\r')) --b;
    return s.substr(a, b - a);
}

bool parse_int(const std::string &v, int &out) {
    char *end = nullptr;
    long x = std::strtol(v.c_str(), &end, 10);
    if (end == v.c_str()) return false;
    out = (int)x;
    return true;
}

int read_v02(FILE *f, const char *path, imp_file_info *info, float *data, int cap) {
    int hdr[4];
    if (std::fread(hdr, sizeof(int), 4, f) != 4)
        return fail(IMP_EFORMAT, "%s: truncated v0.2 header", path);
    info->version = 0.2f;
    info->per_coil = hdr[0];
    info->ncoils = hdr[1];
    info->nslices = hdr[2];
    info->total = hdr[3];
    info->nx = info->ny = info->nz = -1;
    if (hdr[3] < 0 || hdr[0] < 0)
        return fail(IMP_EFORMAT, "%s: negative size in v0.2 header", path);
    if (!data) return IMP_OK;
    if (cap < info->total) return fail(IMP_EINVAL, "%s: buffer holds %d < %d floats", path, cap, info->total);
    if ((int)std::fread(data, sizeof(float), info->total, f) != info->total)
        return fail(IMP_EIO, "%s: short read of %d floats", path, info->total);
    return IMP_OK;
}

int read_v10(FILE *f, const char *path, imp_file_info *info, float *data, int cap) {
    info->version = 1.0f;
    info->per_coil = -1;
    info->ncoils = -1;
    info->nslices = -1;
    info->total = -1;
    info->nx = info->ny = info->nz = -1;
    int binary_size = -1;
    int combined = 0;
    bool saw_eof = false;
    std::string line;
    while (read_line(f, line)) {
        const std::string t = trim(line);
        if (t.empty()) continue;
        if (t.size() >= 2 && t[0] == '/' && t[1] == '/') continue;
        if (t.compare(0, 3, "EOF") == 0) {
            saw_eof = true;
            break;
        }
        if (t == "Binary:") {
            if (binary_size < 0 || info->total < 0)
                return fail(IMP_EFORMAT, "%s: Binary: before file_size/Binary_Size", path);
            if (combined + binary_size > info->total)
                return fail(IMP_EFORMAT, "%s: binary sections exceed file_size", path);
            if (data) {
                if (cap < info->total)
                    return fail(IMP_EINVAL, "%s: buffer holds %d < %d floats", path, cap, info->total);
                if ((int)std::fread(data + combined, sizeof(float), binary_size, f) != binary_size)
                    return fail(IMP_EIO, "%s: short binary section", path);
            } else if (std::fseek(f, (long)binary_size * 4, SEEK_CUR) != 0) {
                return fail(IMP_EIO, "%s: short binary section", path);
            }
            combined += binary_size;
            binary_size = -1;
            continue;
        }
        const size_t eq = t.find('=');
        if (eq == std::string::npos)
            return fail(IMP_EFORMAT, "%s: expected 'key = value', got '%s'", path, t.c_str());
        const std::string key = trim(t.substr(0, eq));
        const std::string val = trim(t.substr(eq + 1));
        int iv = 0;
        if (key == "version") {
            info->version = (float)std::atof(val.c_str());
        } else if (key == "xDimension") {
            if (!parse_int(val, info->nx)) return fail(IMP_EFORMAT, "%s: bad xDimension", path);
        } else if (key == "yDimension") {
            if (!parse_int(val, info->ny)) return fail(IMP_EFORMAT, "%s: bad yDimension", path);
        } else if (key == "zDimension") {
            if (!parse_int(val, info->nz)) return fail(IMP_EFORMAT, "%s: bad zDimension", path);
        } else if (key == "coil_number") {
            if (!parse_int(val, info->ncoils)) return fail(IMP_EFORMAT, "%s: bad coil_number", path);
        } else if (key == "slice_number") {
            if (!parse_int(val, info->nslices)) return fail(IMP_EFORMAT, "%s: bad slice_number", path);
        } else if (key == "file_size") {
            if (!parse_int(val, iv) || iv < 0) return fail(IMP_EFORMAT, "%s: bad file_size", path);
            info->total = iv;
            // coil_number absent: the file is coil-independent (readDataFile_10)
            if (info->ncoils <= 0) info->ncoils = 1;
            info->per_coil = iv / info->ncoils;
        } else if (key == "Binary_Size") {
            if (info->total < 0) return fail(IMP_EFORMAT, "%s: Binary_Size before file_size", path);
            if (!parse_int(val, binary_size) || binary_size < 0)
                return fail(IMP_EFORMAT, "%s: bad Binary_Size", path);
        } else {
            return fail(IMP_EFORMAT, "%s: unknown key '%s'", path, key.c_str());
        }
    }
    if (!saw_eof) return fail(IMP_EFORMAT, "%s: missing EOF marker", path);
    if (info->total < 0) return fail(IMP_EFORMAT, "%s: missing file_size", path);
    if (combined != info->total)
        return fail(IMP_EFORMAT, "%s: binary sections hold %d floats, file_size is %d", path,
                    combined, info->total);
    return IMP_OK;
}

// Matrix Market coordinate file (readMtxFile, kernels/mmio.cpp:668-830, with
// mm_read_banner / mm_read_mtx_crd_size): a "%%MatrixMarket matrix
// coordinate real general" banner (integer values are accepted too), '%'
// comment lines, "rows cols nnz", then nnz lines "i j value", 1-based.
// With cap < 0 only the sizes are read; otherwise up to cap entries are
// stored 0-based.  Indices outside the stated size are an error here (the
// reference checks them only in debug builds).
int read_mtx(const char *path, int *rows, int *cols, int *nnz, int cap, int *I, int *J,
             float *V) {
    File f(path, "r");
    if (!f.f) return fail(IMP_EIO, "%s not found", path);
    std::string line;
    if (!read_line(f.f, line)) return fail(IMP_EFORMAT, "%s: empty file", path);
    std::string banner;
    for (char ch : line) banner.push_back((char)std::tolower((unsigned char)ch));
    char w[5][64] = {};
    if (std::sscanf(banner.c_str(), "%63s %63s %63s %63s %63s", w[0], w[1], w[2], w[3], w[4]) != 5 ||
        std::strcmp(w[0], "%%matrixmarket") != 0 || std::strcmp(w[1], "matrix") != 0)
        return fail(IMP_EFORMAT, "%s: no Matrix Market banner", path);
    if (std::strcmp(w[2], "coordinate") != 0 ||
        (std::strcmp(w[3], "real") != 0 && std::strcmp(w[3], "integer") != 0) ||
        std::strcmp(w[4], "general") != 0)
        return fail(IMP_EFORMAT, "%s: Matrix Market type [%s %s %s] is not supported "
                                 "(coordinate real general)", path, w[2], w[3], w[4]);
    do {
        if (!read_line(f.f, line)) return fail(IMP_EFORMAT, "%s: missing size line", path);
        line = trim(line);
    } while (line.empty() || line[0] == '%');
    int r = 0, c = 0, n = 0;
    if (std::sscanf(line.c_str(), "%d %d %d", &r, &c, &n) != 3 || r <= 0 || c <= 0 || n < 0)
        return fail(IMP_EFORMAT, "%s: bad size line '%s'", path, line.c_str());
    if (rows) *rows = r;
    if (cols) *cols = c;
    if (nnz) *nnz = n;
    if (cap < 0) return IMP_OK;
    if (cap < n) return fail(IMP_EINVAL, "%s: %d entries do not fit in %d", path, n, cap);
    for (int e = 0; e < n; ++e) {
        int ci = 0, cj = 0;
        double v = 0.0;
        if (std::fscanf(f.f, "%d %d %lg", &ci, &cj, &v) != 3)
            return fail(IMP_EFORMAT, "%s: entry %d unreadable", path, e + 1);
        if (ci < 1 || ci > r || cj < 1 || cj > c)
            return fail(IMP_EFORMAT, "%s: entry %d (%d, %d) outside the %d x %d matrix", path,
                        e + 1, ci, cj, r, c);
        I[e] = ci - 1;
        J[e] = cj - 1;
        V[e] = (float)v;
    }
    return IMP_OK;
}

}  // namespace

extern "C" int imp_read_mtx_size(const char *path, int *num_rows, int *num_cols,
                                 int *num_nonzeros) {
    if (!path || !num_rows || !num_cols || !num_nonzeros)
        return fail(IMP_EINVAL, "imp_read_mtx_size: NULL argument");
    return read_mtx(path, num_rows, num_cols, num_nonzeros, -1, nullptr, nullptr, nullptr);
}

extern "C" int imp_read_mtx(const char *path, int cap, int *I, int *J, float *V) {
    if (!path || cap < 0 || (cap > 0 && (!I || !J || !V)))
        return fail(IMP_EINVAL, "imp_read_mtx: bad arguments");
    return read_mtx(path, nullptr, nullptr, nullptr, cap, I, J, V);
}

extern "C" int imp_read_data_file(const char *path, imp_file_info *info, float *data, int cap) {
    if (!path || !info) return fail(IMP_EINVAL, "imp_read_data_file: NULL argument");
    File f(path, "rb");
    if (!f.f) return fail(IMP_EIO, "%s not found", path);
    float v = 0.f;
    if (std::fread(&v, sizeof(float), 1, f.f) != 1) return fail(IMP_EFORMAT, "%s: empty file", path);
    if (v == 0.2f) return read_v02(f.f, path, info, data, cap);
    std::rewind(f.f);
    return read_v10(f.f, path, info, data, cap);
}

extern "C" int imp_write_data_file_v02(const char *path, int per_coil, int ncoils, int nslices,
                                       const float *data, int total) {
    if (!path || (!data && total > 0) || total < 0)
        return fail(IMP_EINVAL, "imp_write_data_file_v02: bad arguments");
    File f(path, "wb");
    if (!f.f) return fail(IMP_EIO, "cannot open %s for writing", path);
    const float v = 0.2f;
    const int hdr[4] = {per_coil, ncoils, nslices, total};
    if (std::fwrite(&v, sizeof(float), 1, f.f) != 1 || std::fwrite(hdr, sizeof(int), 4, f.f) != 4 ||
        (int)std::fwrite(data, sizeof(float), total, f.f) != total)
        return fail(IMP_EIO, "%s: write failed", path);
    return IMP_OK;
}

extern "C" int imp_write_data_file_v10(const char *path, int nx, int ny, int nz, int ncoils,
                                       int nslices, const float *data, int total) {
    if (!path || (!data && total > 0) || total < 0 || ncoils == 0 || ncoils < -1)
        return fail(IMP_EINVAL, "imp_write_data_file_v10: bad arguments");
    const int nsec = ncoils == -1 ? 1 : ncoils;
    if (total % nsec != 0)
        return fail(IMP_EINVAL, "imp_write_data_file_v10: total %d not divisible by %d coils",
                    total, nsec);
    File f(path, "wb");
    if (!f.f) return fail(IMP_EIO, "cannot open %s for writing", path);
    std::fprintf(f.f, "version = %f\n", 1.0);
    std::fprintf(f.f, "xDimension = %d\n", nx);
    std::fprintf(f.f, "yDimension = %d\n", ny);
    std::fprintf(f.f, "zDimension = %d\n", nz);
    if (ncoils != -1) std::fprintf(f.f, "coil_number = %d\n", ncoils);
    std::fprintf(f.f, "slice_number = %d\n", nslices);
    std::fprintf(f.f, "file_size = %d\n", total);
    const int per = total / nsec;
    for (int c = 0; c < nsec; ++c) {
        std::fprintf(f.f, "//Coil No. %d\n", c);
        std::fprintf(f.f, "Binary_Size = %d\n", per);
        std::fprintf(f.f, "Binary:\n");
        if ((int)std::fwrite(data + (size_t)c * per, sizeof(float), per, f.f) != per)
            return fail(IMP_EIO, "%s: write failed", path);
    }
    std::fprintf(f.f, "EOF\n");
    if (std::ferror(f.f)) return fail(IMP_EIO, "%s: write failed", path);
    return IMP_OK;
}
