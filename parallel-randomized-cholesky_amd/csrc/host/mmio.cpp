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
#include <algorithm>
#include <memory>
#include <thread>
#include <vector>

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

int read_mtx(const std::string &path, MtxTriplets &t, std::string &err, int threads)
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
    // Entries are parsed in parallel: the entry text is cut into pieces at
    // line ends, every thread parses its piece into its own arrays, and the
    // pieces are concatenated in file order.  A piece's entries are counted
    // as they are parsed, so the first nz entries of the file are kept.
    const char *body = p;
    const size_t blen = strlen(body);
    int nthreads = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
    nthreads = std::max(1, std::min(nthreads, 64));
    if (blen < ((size_t)1 << 20)) nthreads = 1;      // small files: one piece
    std::vector<size_t> cut((size_t)nthreads + 1, blen);
    cut[0] = 0;
    for (int k = 1; k < nthreads; ++k) {
        size_t c = std::max(cut[(size_t)k - 1], blen * (size_t)k / (size_t)nthreads);
        while (c < blen && body[c - 1] != '\n') ++c;
        cut[(size_t)k] = c;
    }
    struct Piece {
        std::vector<int> r, c;
        std::vector<double> v;
        long long bad_line = -1;       // first malformed line of the piece (0-based within it)
        int bad_kind = 0;              // 1 truncated, 2 missing value, 3 index out of range
    };
    std::vector<Piece> pieces((size_t)nthreads);
    auto parse = [&](int k) {
        Piece &pc = pieces[(size_t)k];
        const char *q = body + cut[(size_t)k];
        const char *qe = body + cut[(size_t)k + 1];
        // an entry takes at least 4 bytes ("1 1\n"), 6 with a value
        const size_t guess = (size_t)(qe - q) / (pattern ? 4 : 6) + 1;
        pc.r.reserve(guess); pc.c.reserve(guess); pc.v.reserve(guess);
        long long line = 0;
        char *end = nullptr;
        while (q < qe) {
            const char *s0 = skip_ws(q);
            if (s0 >= qe) break;
            if (*s0 == '\n') { q = s0 + 1; ++line; continue; }
            long long r = strtoll(s0, &end, 10);
            if (end == s0) { pc.bad_line = line; pc.bad_kind = 1; return; }
            const char *s1 = end;
            long long c = strtoll(s1, &end, 10);
            if (end == s1) { pc.bad_line = line; pc.bad_kind = 1; return; }
            double v = 1.0;
            if (!pattern) {
                const char *s2 = end;
                v = strtod(s2, &end);
                if (end == s2) { pc.bad_line = line; pc.bad_kind = 2; return; }
            }
            if (end > qe) { pc.bad_line = line; pc.bad_kind = 1; return; }   // ran into the next piece
            if (r < 1 || r > nr || c < 1 || c > nc) { pc.bad_line = line; pc.bad_kind = 3; return; }
            pc.r.push_back((int)(r - 1));
            pc.c.push_back((int)(c - 1));
            pc.v.push_back(v);
            const char *nl = (const char *)memchr(end, '\n', (size_t)(qe - end));
            q = nl ? nl + 1 : qe;
            ++line;
        }
    };
    if (nthreads == 1) {
        parse(0);
    } else {
        std::vector<std::thread> th;
        for (int k = 0; k < nthreads; ++k) th.emplace_back(parse, k);
        for (auto &x : th) x.join();
    }
    long long got = 0;
    for (int k = 0; k < nthreads; ++k) {
        const Piece &pc = pieces[(size_t)k];
        if (pc.bad_kind != 0 && got + (long long)pc.r.size() < nz) {
            const long long e = got + (long long)pc.r.size();
            if (pc.bad_kind == 1) err = "truncated entry list at entry " + std::to_string(e);
            else if (pc.bad_kind == 2) err = "missing value at entry " + std::to_string(e);
            else err = "index out of range at entry " + std::to_string(e);
            return RC_ERR_FORMAT;
        }
        got += (long long)pc.r.size();
        if (pc.bad_kind != 0 || got >= nz) break;
    }
    if (got < nz) { err = "truncated entry list at entry " + std::to_string(got); return RC_ERR_FORMAT; }
    const size_t cap = (size_t)nz * (sym ? 2 : 1);
    t.rows.reserve(cap); t.cols.reserve(cap); t.vals.reserve(cap);
    long long left = nz;
    for (int k = 0; k < nthreads && left > 0; ++k) {
        Piece &pc = pieces[(size_t)k];
        const size_t take = (size_t)std::min<long long>(left, (long long)pc.r.size());
        for (size_t e = 0; e < take; ++e) {
            t.rows.push_back(pc.r[e]);
            t.cols.push_back(pc.c[e]);
            t.vals.push_back(pc.v[e]);
            if (sym && pc.r[e] != pc.c[e]) {
                t.rows.push_back(pc.c[e]);
                t.cols.push_back(pc.r[e]);
                t.vals.push_back(sym == 2 ? -pc.v[e] : pc.v[e]);
            }
        }
        left -= (long long)take;
        std::vector<int>().swap(pc.r);
        std::vector<int>().swap(pc.c);
        std::vector<double>().swap(pc.v);
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
