// The reference's on-disk data formats, host side, so that operator pieces and
// preconditioners written by ScalIT (or for it) load into a context and back:
//
// * sequential unformatted (saveData/loadData with binMode = T,
//   src/libutil/serial/io/io.f90:9-33, 98-126): one record per array, the
//   payload framed by a 4-byte little-endian byte count before and after it;
// * list-directed text (the same routines with binMode = F): the values as
//   decimal numbers separated by blanks, commas or line ends, a Fortran
//   exponent letter D accepted and r*c repeat counts expanded on reading;
//   written with 17 significant digits so a value reads back exactly;
// * direct access, one record (saveDataDir/loadDataDir,
//   src/libutil/serial/io/iodirect.f90:10-36, 68-93): RECL is the array's
//   IOLENGTH, so the file is the N doubles and nothing else.
//
// * per-point direct-access records (P2SaveDataGrid,
//   src/libutil/parallel/p_io/p2io_grid.f90:62-83, the OUTH / H0GM files of
//   the ja3 generator): record p (1-based) holds the NG x NG block of grid
//   point p, column-major, RECL = NG*NG doubles.  The operator keeps OUTH
//   with the point index fastest, so reading and writing transpose.
//
// scalit_layout restates calOSB/calDataInfo (src/iterate/module/osb.init.h:2-49):
// where each level's piece starts in the concatenated H0, DEP and VOSB arrays.
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/scalit.h"

namespace {

struct File {
    FILE *f = nullptr;
    File(const char *path, const char *mode) : f(path ? std::fopen(path, mode) : nullptr) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
};

bool write_all(FILE *f, const void *p, size_t bytes) { return std::fwrite(p, 1, bytes, f) == bytes; }
bool read_all(FILE *f, void *p, size_t bytes) { return std::fread(p, 1, bytes, f) == bytes; }

// one list-directed item: "r*c" (r copies of c) or a plain number
bool parse_item(const std::string &tok, long &rep, double &val)
{
    std::string t = tok;
    rep = 1;
    const size_t star = t.find('*');
    if (star != std::string::npos) {
        char *end = nullptr;
        rep = std::strtol(t.c_str(), &end, 10);
        if (end != t.c_str() + star || rep < 1) return false;
        t = t.substr(star + 1);
    }
    for (char &ch : t)
        if (ch == 'd' || ch == 'D') ch = 'e';
    char *end = nullptr;
    errno = 0;
    val = std::strtod(t.c_str(), &end);
    return end == t.c_str() + t.size() && !t.empty() && errno != ERANGE;
}

}  // namespace

extern "C" {

int scalit_save_data(const char *path, long n, const double *data, int bin_mode)
{
    if (!path || n < 0 || (n > 0 && !data)) return SCALIT_EINVAL;
    File f(path, bin_mode ? "wb" : "w");
    if (!f.f) return SCALIT_EIO;
    if (bin_mode) {
        const uint64_t bytes = (uint64_t)n * sizeof(double);
        if (bytes > 0x7fffffffULL) return SCALIT_EINVAL;   // one record, 4-byte markers
        const int32_t m = (int32_t)bytes;
        if (!write_all(f.f, &m, 4) || !write_all(f.f, data, bytes) || !write_all(f.f, &m, 4)) return SCALIT_EIO;
    } else {
        for (long i = 0; i < n; i++)
            if (std::fprintf(f.f, (i % 4 == 3 || i == n - 1) ? " %.17E\n" : " %.17E", data[i]) < 0) return SCALIT_EIO;
    }
    return SCALIT_OK;
}

int scalit_load_data(const char *path, long n, double *data, int bin_mode)
{
    if (!path || n < 0 || (n > 0 && !data)) return SCALIT_EINVAL;
    File f(path, bin_mode ? "rb" : "r");
    if (!f.f) return SCALIT_EIO;
    if (bin_mode) {
        // READ of n values from the first record: the record must hold at least n
        int32_t m0 = 0, m1 = 0;
        if (!read_all(f.f, &m0, 4) || m0 < 0 || (uint64_t)m0 < (uint64_t)n * sizeof(double)) return SCALIT_EIO;
        if (!read_all(f.f, data, (size_t)n * sizeof(double))) return SCALIT_EIO;
        if (std::fseek(f.f, (long)m0 - (long)(n * sizeof(double)), SEEK_CUR) != 0 || !read_all(f.f, &m1, 4) || m1 != m0)
            return SCALIT_EIO;
        return SCALIT_OK;
    }
    long got = 0;
    std::string tok;
    auto flush = [&]() -> bool {
        if (tok.empty()) return true;
        long rep;
        double v;
        if (!parse_item(tok, rep, v)) return false;
        for (long k = 0; k < rep && got < n; k++) data[got++] = v;
        tok.clear();
        return true;
    };
    int ch;
    while (got < n && (ch = std::fgetc(f.f)) != EOF) {
        if (ch == ' ' || ch == ',' || ch == '\n' || ch == '\r' || ch == '\t') {
            if (!flush()) return SCALIT_EIO;
        } else {
            tok.push_back((char)ch);
        }
    }
    if (!flush()) return SCALIT_EIO;
    return got == n ? SCALIT_OK : SCALIT_EIO;
}

int scalit_save_data_dir(const char *path, long n, const double *data)
{
    if (!path || n < 0 || (n > 0 && !data)) return SCALIT_EINVAL;
    File f(path, "wb");
    if (!f.f) return SCALIT_EIO;
    return write_all(f.f, data, (size_t)n * sizeof(double)) ? SCALIT_OK : SCALIT_EIO;
}

int scalit_load_data_dir(const char *path, long n, double *data)
{
    if (!path || n < 0 || (n > 0 && !data)) return SCALIT_EINVAL;
    File f(path, "rb");
    if (!f.f) return SCALIT_EIO;
    return read_all(f.f, data, (size_t)n * sizeof(double)) ? SCALIT_OK : SCALIT_EIO;
}

int scalit_load_data_grid(const char *path, long npts, int ng, double *outh)
{
    if (!path || npts < 1 || ng < 1 || !outh) return SCALIT_EINVAL;
    File in(path, "rb");
    if (!in.f) return SCALIT_EIO;
    const long nn = (long)ng * ng;
    std::vector<double> rec(nn);
    for (long p = 0; p < npts; p++) {
        if (!read_all(in.f, rec.data(), sizeof(double) * nn)) return SCALIT_EIO;
        for (long e = 0; e < nn; e++) outh[p + npts * e] = rec[e];
    }
    return SCALIT_OK;
}

int scalit_save_data_grid(const char *path, long npts, int ng, const double *outh)
{
    if (!path || npts < 1 || ng < 1 || !outh) return SCALIT_EINVAL;
    File out(path, "wb");
    if (!out.f) return SCALIT_EIO;
    const long nn = (long)ng * ng;
    std::vector<double> rec(nn);
    for (long p = 0; p < npts; p++) {
        for (long e = 0; e < nn; e++) rec[e] = outh[p + npts * e];
        if (!write_all(out.f, rec.data(), sizeof(double) * nn)) return SCALIT_EIO;
    }
    return SCALIT_OK;
}

int scalit_layout(int sF, const int *sN, const int *sDEP, int ndvr, long *h0_start, long *dep_start,
                  long *vosb_start, long *totals)
{
    if (sF < 1 || sF > SCALIT_FMAX || !sN || !sDEP || !totals) return SCALIT_EINVAL;
    if (ndvr && sF < 2) return SCALIT_EINVAL;
    long dim[SCALIT_FMAX], blk[SCALIT_FMAX];
    dim[0] = 1;
    for (int i = 0; i < sF; i++) {
        if (sN[i] < 1) return SCALIT_EINVAL;
        if (i + 1 < sF) dim[i + 1] = dim[i] * sN[i];
    }
    blk[sF - 1] = 1;
    for (int i = sF - 1; i >= 1; i--) blk[i - 1] = blk[i] * sN[i];
    const long len = blk[0] * sN[0];
    long h0 = 0, dep = 0, vosb = 0;
    for (int i = 0; i < sF; i++) {
        if (h0_start) h0_start[i] = h0;
        if (dep_start) dep_start[i] = dep;
        if (vosb_start) vosb_start[i] = vosb;
        h0 += (ndvr && i == sF - 1) ? 0 : (long)sN[i] * sN[i];
        dep += sDEP[i] ? dim[i] : 0;
        vosb += blk[i] * sN[i] * sN[i];
    }
    totals[0] = len;                          // myLen
    totals[1] = h0;                           // myH0%mLen
    totals[2] = dep;                          // myDEP%mLen
    totals[3] = vosb;                         // myVOSB%mLen
    totals[4] = ndvr ? (long)sN[sF - 1] * len : 0;   // outLen
    return SCALIT_OK;
}

}  // extern "C"
