// mmio.cpp -- Matrix Market coordinate reader/writer (the reference uses the
// un-vendored fast_matrix_market: pre_process.hpp:333-334 reads triplets,
// driver_local.cpp:969-983 writes the factor).  Own implementation:
//   * `matrix coordinate {real|integer|pattern} {general|symmetric|skew-symmetric}`;
//   * symmetric / skew-symmetric files are expanded to both triangles (the
//     off-diagonal mirror is added once, the diagonal is not duplicated);
//   * pattern entries read as 1.0;
//   * indices are 1-based in the file, 0-based in memory.
#include "mmio.hpp"

#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "../../../include/rchol.h"

namespace rc {

namespace {

struct FileCloser {
    void operator()(FILE *f) const { if (f) fclose(f); }
};

bool ieq(const char *a, const char *b)
{
    for (; *a && *b; ++a, ++b)
        if (std::tolower((unsigned char)*a) != std::tolower((unsigned char)*b)) return false;
    return *a == *b;
}

// Fast unsigned/signed integer and double parsing from a cursor.
inline const char *skip_ws(const char *p)
{
    while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
    return p;
}

}  // namespace

int read_mtx(const std::string &path, MtxTriplets &t, std::string &err)
{
    std::unique_ptr<FILE, FileCloser> f(fopen(path.c_str(), "rb"));
    if (!f) { err = "cannot open " + path + ": " + strerror(errno); return RC_ERR_IO; }
    // read whole file (streams of several GB are read in one pass)
    if (fseek(f.get(), 0, SEEK_END) != 0) { err = "cannot seek " + path; return RC_ERR_IO; }
    long sz = ftell(f.get());
    if (sz < 0) { err = "cannot size " + path; return RC_ERR_IO; }
    rewind(f.get());
    std::string buf((size_t)sz, '\0');
    if (sz > 0 && fread(&buf[0], 1, (size_t)sz, f.get()) != (size_t)sz) { err = "short read " + path; return RC_ERR_IO; }
    buf.push_back('\0');
    const char *p = buf.c_str();

    // header
    char w[5][64] = {{0}};
    if (sscanf(p, "%63s %63s %63s %63s %63s", w[0], w[1], w[2], w[3], w[4]) != 5 || !ieq(w[0], "%%MatrixMarket")) {
        err = "missing %%MatrixMarket header";
        return RC_ERR_FORMAT;
    }
    if (!ieq(w[1], "matrix") || !ieq(w[2], "coordinate")) { err = "only 'matrix coordinate' is supported"; return RC_ERR_FORMAT; }
    bool pattern = ieq(w[3], "pattern");
    if (!pattern && !ieq(w[3], "real") && !ieq(w[3], "integer") && !ieq(w[3], "double")) {
        err = std::string("unsupported field ") + w[3];
        return RC_ERR_FORMAT;
    }
    int sym = 0;  // 0 general, 1 symmetric, 2 skew
    if (ieq(w[4], "symmetric")) sym = 1;
    else if (ieq(w[4], "skew-symmetric")) sym = 2;
    else if (!ieq(w[4], "general")) { err = std::string("unsupported symmetry ") + w[4]; return RC_ERR_FORMAT; }
    t.symmetric = sym != 0;

    // skip header + comment lines
    while (*p) {
        const char *q = skip_ws(p);
        if (*q == '%' || *q == '\n') {
            const char *nl = strchr(p, '\n');
            if (!nl) { p += strlen(p); break; }
            p = nl + 1;
            continue;
        }
        break;
    }
    char *end = nullptr;
    long long nr = strtoll(p, &end, 10); p = end;
    long long nc = strtoll(p, &end, 10); p = end;
    long long nz = strtoll(p, &end, 10);
    if (end == p || nr < 0 || nc < 0 || nz < 0) { err = "bad size line"; return RC_ERR_FORMAT; }
    p = end;
    if (nr > 2147483647LL || nc > 2147483647LL) { err = "dimension exceeds int32"; return RC_ERR_FORMAT; }
    t.nrows = (int)nr;
    t.ncols = (int)nc;
    t.rows.clear(); t.cols.clear(); t.vals.clear();
    size_t cap = (size_t)nz * (sym ? 2 : 1);
    t.rows.reserve(cap); t.cols.reserve(cap); t.vals.reserve(cap);
    for (long long e = 0; e < nz; ++e) {
        long long r = strtoll(p, &end, 10);
        if (end == p) { err = "truncated entry list at entry " + std::to_string(e); return RC_ERR_FORMAT; }
        p = end;
        long long c = strtoll(p, &end, 10);
        if (end == p) { err = "truncated entry list at entry " + std::to_string(e); return RC_ERR_FORMAT; }
        p = end;
        double v = 1.0;
        if (!pattern) {
            v = strtod(p, &end);
            if (end == p) { err = "missing value at entry " + std::to_string(e); return RC_ERR_FORMAT; }
            p = end;
        }
        if (r < 1 || r > nr || c < 1 || c > nc) { err = "index out of range at entry " + std::to_string(e); return RC_ERR_FORMAT; }
        t.rows.push_back((int)(r - 1));
        t.cols.push_back((int)(c - 1));
        t.vals.push_back(v);
        if (sym && r != c) {
            t.rows.push_back((int)(c - 1));
            t.cols.push_back((int)(r - 1));
            t.vals.push_back(sym == 2 ? -v : v);
        }
    }
    return RC_OK;
}

int write_mtx_csc(const std::string &path, int nrows, int ncols, const int64_t *colptr, const int *rowind,
                  const double *val, std::string &err)
{
    std::unique_ptr<FILE, FileCloser> f(fopen(path.c_str(), "wb"));
    if (!f) { err = "cannot open " + path + " for writing: " + strerror(errno); return RC_ERR_IO; }
    fprintf(f.get(), "%%%%MatrixMarket matrix coordinate real general\n");
    fprintf(f.get(), "%d %d %lld\n", nrows, ncols, (long long)colptr[ncols]);
    for (int j = 0; j < ncols; ++j)
        for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p)
            fprintf(f.get(), "%d %d %.17g\n", rowind[p] + 1, j + 1, val[p]);
    if (fflush(f.get()) != 0) { err = "write failed " + path; return RC_ERR_IO; }
    return RC_OK;
}

}  // namespace rc
