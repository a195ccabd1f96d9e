// sparse.cpp -- host preprocessing on the input boundary (see sparse.hpp).
#include "sparse.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <queue>
#include <utility>

#include "../../../include/rchol.h"
#include "philox.h"

namespace rc {

namespace {

void sort_columns(Csc &a)
{
    std::vector<std::pair<int, double>> tmp;
    for (int j = 0; j < a.n; ++j) {
        int64_t b = a.ptr[j], e = a.ptr[j + 1];
        bool sorted = true;
        for (int64_t p = b + 1; p < e; ++p)
            if (a.ind[p - 1] > a.ind[p] || (a.ind[p - 1] == a.ind[p] && a.val[p - 1] > a.val[p])) { sorted = false; break; }
        if (sorted) continue;
        tmp.resize((size_t)(e - b));
        for (int64_t p = b; p < e; ++p) tmp[(size_t)(p - b)] = {a.ind[p], a.val[p]};
        std::sort(tmp.begin(), tmp.end());
        for (int64_t p = b; p < e; ++p) { a.ind[p] = tmp[(size_t)(p - b)].first; a.val[p] = tmp[(size_t)(p - b)].second; }
    }
}

}  // namespace

int build_laplacian(int n, int64_t m, const int *rows, const int *cols, const double *vals, Csc &out)
{
    if (n < 0 || m < 0) return RC_ERR_ARG;
    bool add_diag = true;
    for (int64_t t = 0; t < m; ++t) {
        if (rows[t] < 0 || rows[t] >= n || cols[t] < 0 || cols[t] >= n) return RC_ERR_FORMAT;
        if (rows[t] == cols[t]) add_diag = false;              // pre_process.hpp:338-345
    }
    std::vector<double> dsum(add_diag ? (size_t)n : 0, 0.0);
    std::vector<int64_t> cnt((size_t)n + 1, 0);
    for (int64_t t = 0; t < m; ++t) {
        if (rows[t] != cols[t]) {
            if (vals[t] == 0.0) continue;
            if (add_diag) dsum[(size_t)cols[t]] += std::fabs(vals[t]);   // :217-228
        }
        cnt[(size_t)cols[t] + 1]++;
    }
    if (add_diag)
        for (int i = 0; i < n; ++i) cnt[(size_t)i + 1]++;
    out.n = n;
    out.ptr.assign((size_t)n + 1, 0);
    for (int i = 0; i < n; ++i) out.ptr[(size_t)i + 1] = out.ptr[(size_t)i] + cnt[(size_t)i + 1];
    out.ind.assign((size_t)out.ptr[(size_t)n], 0);
    out.val.assign((size_t)out.ptr[(size_t)n], 0.0);
    std::vector<int64_t> pos(out.ptr.begin(), out.ptr.end() - 1);
    for (int64_t t = 0; t < m; ++t) {
        double v;
        if (rows[t] != cols[t]) {
            if (vals[t] == 0.0) continue;
            v = -std::fabs(vals[t]);                                     // :197-208
        } else {
            v = std::fabs(vals[t]);
        }
        int64_t p = pos[(size_t)cols[t]]++;
        out.ind[(size_t)p] = rows[t];
        out.val[(size_t)p] = v;
    }
    if (add_diag)
        for (int i = 0; i < n; ++i) {
            int64_t p = pos[(size_t)i]++;
            out.ind[(size_t)p] = i;
            out.val[(size_t)p] = dsum[(size_t)i];
        }
    sort_columns(out);                                                   // :240-262
    return RC_OK;
}

bool is_symmetric(const Csc &a)
{
    // transpose by counting and compare column by column (rows are sorted)
    std::vector<int64_t> tp((size_t)a.n + 1, 0);
    for (int64_t p = 0; p < a.nnz(); ++p) tp[(size_t)a.ind[(size_t)p] + 1]++;
    for (int i = 0; i < a.n; ++i) tp[(size_t)i + 1] += tp[(size_t)i];
    if (tp != a.ptr) return false;
    std::vector<int64_t> pos(tp.begin(), tp.end() - 1);
    std::vector<int> ti((size_t)a.nnz());
    std::vector<double> tv((size_t)a.nnz());
    for (int j = 0; j < a.n; ++j)
        for (int64_t p = a.ptr[(size_t)j]; p < a.ptr[(size_t)j + 1]; ++p) {
            int64_t q = pos[(size_t)a.ind[(size_t)p]]++;
            ti[(size_t)q] = j;
            tv[(size_t)q] = a.val[(size_t)p];
        }
    // columns of the transpose come out sorted by row; duplicates may differ
    // in value order, so compare as multisets per column
    for (int j = 0; j < a.n; ++j) {
        int64_t b = a.ptr[(size_t)j], e = a.ptr[(size_t)j + 1];
        for (int64_t p = b; p < e; ++p) {
            if (ti[(size_t)p] != a.ind[(size_t)p]) return false;
        }
        bool same = true;
        for (int64_t p = b; p < e; ++p)
            if (tv[(size_t)p] != a.val[(size_t)p]) { same = false; break; }
        if (!same) {
            std::vector<double> x(a.val.begin() + b, a.val.begin() + e), y(tv.begin() + b, tv.begin() + e);
            std::sort(x.begin(), x.end());
            std::sort(y.begin(), y.end());
            if (x != y) return false;
        }
    }
    return true;
}

int ground_append(const Csc &a, Csc &out)
{
    const int n = a.n;
    double total = 0.0;
    std::vector<double> app((size_t)n);
    for (int j = 0; j < n; ++j) {
        double cs = 0.0;
        for (int64_t p = a.ptr[(size_t)j]; p < a.ptr[(size_t)j + 1]; ++p) cs += a.val[(size_t)p];
        total += cs;
        double v = -cs;
        app[(size_t)j] = v > 0.0 ? 0.0 : v;                              // write_graph.jl:375-377
    }
    if (total < 0.0 && std::fabs(total) > 1e-4) return RC_ERR_NOT_DOMINANT;   // :369-372
    if (!(total > 1e-9)) { out = a; return 0; }                                // :373
    double corner = 0.0;
    int64_t extra = 0;
    for (int j = 0; j < n; ++j) {
        corner += app[(size_t)j];
        if (app[(size_t)j] != 0.0) ++extra;
    }
    corner = -corner;                                                           // :381
    out.n = n + 1;
    out.ptr.assign((size_t)n + 2, 0);
    out.ind.resize((size_t)(a.nnz() + 2 * extra + 1));
    out.val.resize(out.ind.size());
    int64_t q = 0;
    for (int j = 0; j < n; ++j) {
        for (int64_t p = a.ptr[(size_t)j]; p < a.ptr[(size_t)j + 1]; ++p) {
            out.ind[(size_t)q] = a.ind[(size_t)p];
            out.val[(size_t)q] = a.val[(size_t)p];
            ++q;
        }
        if (app[(size_t)j] != 0.0) { out.ind[(size_t)q] = n; out.val[(size_t)q] = app[(size_t)j]; ++q; }
        out.ptr[(size_t)j + 1] = q;
    }
    for (int j = 0; j < n; ++j)
        if (app[(size_t)j] != 0.0) { out.ind[(size_t)q] = j; out.val[(size_t)q] = app[(size_t)j]; ++q; }
    out.ind[(size_t)q] = n;
    out.val[(size_t)q] = corner;
    ++q;
    out.ptr[(size_t)n + 1] = q;
    return 1;
}

void trim_last(const Csc &a, Csc &out)
{
    const int n = a.n;
    Csc t;
    t.n = n - 1;
    t.ptr.assign((size_t)n, 0);
    t.ind.reserve((size_t)a.nnz());
    t.val.reserve((size_t)a.nnz());
    for (int j = 0; j < n - 1; ++j) {
        for (int64_t p = a.ptr[(size_t)j]; p < a.ptr[(size_t)j + 1]; ++p)
            if (a.ind[(size_t)p] != n - 1) { t.ind.push_back(a.ind[(size_t)p]); t.val.push_back(a.val[(size_t)p]); }
        t.ptr[(size_t)j + 1] = (int64_t)t.ind.size();
    }
    out = std::move(t);
}

void permute_sym(const Csc &a, const std::vector<int> &perm, Csc &out)
{
    const int n = a.n;
    std::vector<int> inv((size_t)n);
    for (int i = 0; i < n; ++i) inv[(size_t)perm[(size_t)i]] = i;
    Csc t;
    t.n = n;
    t.ptr.assign((size_t)n + 1, 0);
    for (int j = 0; j < n; ++j) {
        int oc = perm[(size_t)j];
        t.ptr[(size_t)j + 1] = t.ptr[(size_t)j] + (a.ptr[(size_t)oc + 1] - a.ptr[(size_t)oc]);
    }
    t.ind.resize((size_t)t.ptr[(size_t)n]);
    t.val.resize((size_t)t.ptr[(size_t)n]);
    for (int j = 0; j < n; ++j) {
        int oc = perm[(size_t)j];
        int64_t q = t.ptr[(size_t)j];
        for (int64_t p = a.ptr[(size_t)oc]; p < a.ptr[(size_t)oc + 1]; ++p, ++q) {
            t.ind[(size_t)q] = inv[(size_t)a.ind[(size_t)p]];
            t.val[(size_t)q] = a.val[(size_t)p];
        }
    }
    sort_columns(t);
    out = std::move(t);
}

std::vector<int> random_perm(int n, uint64_t seed)
{
    std::vector<int> p((size_t)n);
    std::iota(p.begin(), p.end(), 0);
    for (int i = n - 1; i > 0; --i) {                   // Fisher-Yates over Philox
        U4 o = philox_keyed(seed, (uint32_t)i, 0u, kTagPerm);
        uint64_t r = ((uint64_t)o.y << 32) | o.x;
        int j = (int)(((unsigned __int128)r * (unsigned __int128)(uint64_t)(i + 1)) >> 64);
        std::swap(p[(size_t)i], p[(size_t)j]);
    }
    return p;
}

std::vector<int> nnz_sort_perm(const Csc &a, int nvert, uint64_t seed, int count_skip)
{
    std::vector<int> p1 = random_perm(nvert, seed);     // randperm (write_graph.jl:333-334)
    auto count = [&](int v) {
        int64_t c = a.ptr[(size_t)v + 1] - a.ptr[(size_t)v];
        if (count_skip >= 0)
            for (int64_t p = a.ptr[(size_t)v]; p < a.ptr[(size_t)v + 1]; ++p)
                if (a.ind[(size_t)p] == count_skip) --c;
        return c;
    };
    std::vector<int64_t> c((size_t)nvert);
    for (int i = 0; i < nvert; ++i) c[(size_t)i] = count(p1[(size_t)i]);
    std::vector<int> idx((size_t)nvert);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return c[(size_t)x] < c[(size_t)y]; });  // sortperm
    std::vector<int> perm((size_t)nvert);
    for (int i = 0; i < nvert; ++i) perm[(size_t)i] = p1[(size_t)idx[(size_t)i]];
    return perm;
}

// Approximate minimum degree on the quotient graph (Amestoy, Davis, Duff,
// SIAM J. Matrix Anal. Appl. 17(4), 1996): variables keep their remaining
// variable neighbours A_i and adjacent elements E_i; eliminating the variable
// p of least approximate degree forms the element L_p = A_p + union of the
// elements in E_p (which are absorbed), and every i in L_p gets the bound
//   d_i = min(n - k - 1, d_i + |L_p \ i|, |A_i \ i| + |L_p \ i| + sum_{e in E_i, e != p} |L_e \ L_p|)
// where |L_e \ L_p| comes from one scan over L_p.  Ties are broken by the
// smaller vertex index, so the ordering is deterministic.  Supervariable
// detection and aggressive absorption are not implemented.
std::vector<int> amd_perm(const Csc &a, int nvert)
{
    const int n = nvert;
    std::vector<std::vector<int>> A((size_t)n), E((size_t)n), Le((size_t)n);
    for (int j = 0; j < n; ++j)
        for (int64_t p = a.ptr[(size_t)j]; p < a.ptr[(size_t)j + 1]; ++p) {
            int i = a.ind[(size_t)p];
            if (i != j && i < n) A[(size_t)j].push_back(i);
        }
    for (auto &v : A) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    std::vector<char> elem((size_t)n, 0), dead((size_t)n, 0);
    std::vector<int> deg((size_t)n), mark((size_t)n, -1), wmark((size_t)n, -1), w((size_t)n, 0);
    using QE = std::pair<int, int>;
    std::priority_queue<QE, std::vector<QE>, std::greater<QE>> pq;
    for (int i = 0; i < n; ++i) {
        deg[(size_t)i] = (int)A[(size_t)i].size();
        pq.push({deg[(size_t)i], i});
    }
    std::vector<int> order;
    order.reserve((size_t)n);
    std::vector<int> Lp;
    for (int k = 0; k < n; ++k) {
        int p = -1;
        while (!pq.empty()) {
            QE t = pq.top();
            pq.pop();
            if (!elem[(size_t)t.second] && t.first == deg[(size_t)t.second]) { p = t.second; break; }
        }
        if (p < 0) break;
        // element L_p
        Lp.clear();
        mark[(size_t)p] = k;
        for (int v : A[(size_t)p])
            if (!elem[(size_t)v] && mark[(size_t)v] != k) { mark[(size_t)v] = k; Lp.push_back(v); }
        for (int e : E[(size_t)p]) {
            if (dead[(size_t)e]) continue;
            for (int v : Le[(size_t)e])
                if (!elem[(size_t)v] && mark[(size_t)v] != k) { mark[(size_t)v] = k; Lp.push_back(v); }
            dead[(size_t)e] = 1;                      // absorbed into p
            std::vector<int>().swap(Le[(size_t)e]);
        }
        order.push_back(p);
        elem[(size_t)p] = 1;
        std::vector<int>().swap(A[(size_t)p]);
        std::vector<int>().swap(E[(size_t)p]);
        Le[(size_t)p] = Lp;
        // |L_e \ L_p| for every live element adjacent to L_p
        for (int i : Lp) {
            auto &Ei = E[(size_t)i];
            size_t q = 0;
            for (size_t t = 0; t < Ei.size(); ++t) {
                int e = Ei[t];
                if (dead[(size_t)e]) continue;
                Ei[q++] = e;
                if (wmark[(size_t)e] != k) {
                    auto &L = Le[(size_t)e];
                    size_t r = 0;
                    for (size_t u = 0; u < L.size(); ++u)
                        if (!elem[(size_t)L[u]]) L[r++] = L[u];
                    L.resize(r);
                    wmark[(size_t)e] = k;
                    w[(size_t)e] = (int)L.size();
                }
                w[(size_t)e] -= 1;
            }
            Ei.resize(q);
        }
        const int lp = (int)Lp.size();
        for (int i : Lp) {
            auto &Ai = A[(size_t)i];
            size_t q = 0;
            for (size_t t = 0; t < Ai.size(); ++t) {
                int v = Ai[t];
                if (v == p || elem[(size_t)v] || mark[(size_t)v] == k) continue;
                Ai[q++] = v;
            }
            Ai.resize(q);
            long long ext = (long long)Ai.size() + (lp - 1);
            for (int e : E[(size_t)i]) ext += w[(size_t)e];
            E[(size_t)i].push_back(p);
            long long d = std::min<long long>((long long)(n - k - 1), (long long)deg[(size_t)i] + (lp - 1));
            d = std::min(d, ext);
            deg[(size_t)i] = (int)std::max<long long>(d, 0);
            pq.push({deg[(size_t)i], i});
        }
    }
    if ((int)order.size() != n) return {};
    return order;
}

int lower_part(const Csc &a, std::vector<int> &ptr, std::vector<int> &row, std::vector<double> &w)
{
    int64_t cnt = 0;
    for (int j = 0; j < a.n; ++j)
        for (int64_t p = a.ptr[(size_t)j]; p < a.ptr[(size_t)j + 1]; ++p)
            if (a.ind[(size_t)p] > j) ++cnt;
    if (cnt > 2147483647LL) return RC_ERR_ARG;
    ptr.assign((size_t)a.n + 1, 0);
    row.resize((size_t)cnt);
    w.resize((size_t)cnt);
    int q = 0;
    for (int j = 0; j < a.n; ++j) {
        for (int64_t p = a.ptr[(size_t)j]; p < a.ptr[(size_t)j + 1]; ++p)
            if (a.ind[(size_t)p] > j) { row[(size_t)q] = a.ind[(size_t)p]; w[(size_t)q] = std::fabs(a.val[(size_t)p]); ++q; }
        ptr[(size_t)j + 1] = q;
    }
    return RC_OK;
}

int components(const Csc &a, std::vector<int> &comp)
{
    comp.assign((size_t)a.n, -1);
    int nc = 0;
    std::vector<int> stack;
    for (int s = 0; s < a.n; ++s) {
        if (comp[(size_t)s] >= 0) continue;
        comp[(size_t)s] = nc;
        stack.push_back(s);
        while (!stack.empty()) {
            int v = stack.back();
            stack.pop_back();
            for (int64_t p = a.ptr[(size_t)v]; p < a.ptr[(size_t)v + 1]; ++p) {
                int u = a.ind[(size_t)p];
                if (comp[(size_t)u] < 0) { comp[(size_t)u] = nc; stack.push_back(u); }
            }
        }
        ++nc;
    }
    return nc;
}

}  // namespace rc
