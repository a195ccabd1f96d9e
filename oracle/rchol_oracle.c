/*
 * rchol_oracle.c -- CPU oracle (test infrastructure only; see rchol_oracle.h).
 *
 * Single-threaded C restatement of the reference's randomized Cholesky factor
 * and PCG.  Every function cites the reference lines it follows.  Columns are
 * eliminated in index order 0..n-1: every dependency in the reference's DAG
 * runs from a lower to a higher index (a fill edge (max, min) is created by a
 * column k < min), so index order is a valid schedule, and with the canonical
 * merge/sort orders the result does not depend on the schedule at all.
 */
#include "rchol_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Philox4x32-10, from the round function published in Salmon et al. (SC'11) */
/* ------------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += PHILOX_W0; k1 += PHILOX_W1; }
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void philox_keyed(uint64_t seed, uint32_t a, uint32_t b, uint32_t tag, uint32_t out[4])
{
    uint32_t ctr[4] = {a, b, tag, 0u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32(ctr, key, out);
}

static double u53_open0(uint32_t lo, uint32_t hi)
{   /* (0, 1] */
    uint64_t bits = (((uint64_t)hi << 32) | lo) >> 11;
    return (double)(bits + 1) * 0x1.0p-53;
}

static double u53_open1(uint32_t lo, uint32_t hi)
{   /* [0, 1) */
    uint64_t bits = (((uint64_t)hi << 32) | lo) >> 11;
    return (double)bits * 0x1.0p-53;
}

double orc_uniform(uint64_t seed, uint32_t a, uint32_t b, uint32_t tag)
{
    uint32_t o[4];
    philox_keyed(seed, a, b, tag, o);
    return u53_open0(o[0], o[1]);
}

double orc_normal(uint64_t seed, uint32_t index, uint32_t stream)
{
    uint32_t o[4];
    philox_keyed(seed, index, stream, ORC_TAG_NORMAL, o);
    double u1 = u53_open0(o[0], o[1]);
    double u2 = u53_open1(o[2], o[3]);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

void orc_normal_vector(uint64_t seed, int64_t n, uint32_t stream, double *out)
{
    for (int64_t i = 0; i < n; ++i) out[i] = orc_normal(seed, (uint32_t)i, stream);
}

/* ------------------------------------------------------------------------ */
/* Small sort helpers (insertion sort for short runs, qsort beyond).        */
/* ------------------------------------------------------------------------ */
typedef struct { int row; int mult; double w; } entry_t;

static int cmp_row_value(const void *pa, const void *pb)
{
    const entry_t *a = (const entry_t *)pa, *b = (const entry_t *)pb;
    if (a->row != b->row) return a->row < b->row ? -1 : 1;
    if (a->w != b->w) return a->w < b->w ? -1 : 1;
    return 0;
}

static int cmp_value_row(const void *pa, const void *pb)
{
    const entry_t *a = (const entry_t *)pa, *b = (const entry_t *)pb;
    if (a->w != b->w) return a->w < b->w ? -1 : 1;
    if (a->row != b->row) return a->row < b->row ? -1 : 1;
    return 0;
}

typedef struct { int row; double v; } rv_t;

static int cmp_rv(const void *pa, const void *pb)
{
    const rv_t *a = (const rv_t *)pa, *b = (const rv_t *)pb;
    if (a->row != b->row) return a->row < b->row ? -1 : 1;
    if (a->v != b->v) return a->v < b->v ? -1 : 1;
    return 0;
}

/* Sort the entries of every CSC column by row (stable w.r.t. equal rows is
 * not needed: equal (row) entries keep qsort order, callers never rely on it
 * except through values that are summed canonically later). */
static void sort_columns(int n, const int64_t *colptr, int *rowind, double *val)
{
    int64_t maxlen = 0;
    for (int j = 0; j < n; ++j) {
        int64_t len = colptr[j + 1] - colptr[j];
        if (len > maxlen) maxlen = len;
    }
    rv_t *tmp = (rv_t *)malloc((size_t)(maxlen > 0 ? maxlen : 1) * sizeof(rv_t));
    for (int j = 0; j < n; ++j) {
        int64_t b = colptr[j], len = colptr[j + 1] - b;
        for (int64_t t = 0; t < len; ++t) { tmp[t].row = rowind[b + t]; tmp[t].v = val[b + t]; }
        /* (row, value) total order so duplicates land deterministically */
        if (len > 32) qsort(tmp, (size_t)len, sizeof(rv_t), cmp_rv);
        else for (int64_t t = 1; t < len; ++t) {
            rv_t x = tmp[t];
            int64_t s = t - 1;
            while (s >= 0 && (tmp[s].row > x.row || (tmp[s].row == x.row && tmp[s].v > x.v))) {
                tmp[s + 1] = tmp[s];
                --s;
            }
            tmp[s + 1] = x;
        }
        for (int64_t t = 0; t < len; ++t) { rowind[b + t] = tmp[t].row; val[b + t] = tmp[t].v; }
    }
    free(tmp);
}

/* ------------------------------------------------------------------------ */
/* Laplacian construction: gpu_implementation/pre_process.hpp:189-306,338-357 */
/* ------------------------------------------------------------------------ */
int64_t orc_laplacian_from_triplets(int n, int64_t m, const int *rows, const int *cols,
                                    const double *vals, int64_t *colptr, int *rowind,
                                    double *val)
{
    int add_diag = 1;
    for (int64_t t = 0; t < m; ++t) {
        if (rows[t] < 0 || rows[t] >= n || cols[t] < 0 || cols[t] >= n) return -1;
        if (rows[t] == cols[t]) add_diag = 0;          /* pre_process.hpp:338-345 */
    }
    double *dsum = (double *)calloc((size_t)n, sizeof(double));
    int64_t *cnt = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t t = 0; t < m; ++t) {
        if (rows[t] != cols[t]) {
            if (vals[t] == 0.0) continue;              /* explicit zero dropped */
            if (add_diag) dsum[cols[t]] += fabs(vals[t]);   /* :217-228 */
        }
        cnt[cols[t] + 1]++;
    }
    if (add_diag)
        for (int i = 0; i < n; ++i) cnt[i + 1]++;
    colptr[0] = 0;
    for (int i = 0; i < n; ++i) colptr[i + 1] = colptr[i] + cnt[i + 1];
    int64_t *pos = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    for (int i = 0; i < n; ++i) pos[i] = colptr[i];
    for (int64_t t = 0; t < m; ++t) {
        double v;
        if (rows[t] != cols[t]) {
            if (vals[t] == 0.0) continue;
            v = -fabs(vals[t]);                        /* :197-208 */
        } else {
            v = fabs(vals[t]);
        }
        int64_t p = pos[cols[t]]++;
        rowind[p] = rows[t];
        val[p] = v;
    }
    if (add_diag)
        for (int i = 0; i < n; ++i) { int64_t p = pos[i]++; rowind[p] = i; val[p] = dsum[i]; }
    sort_columns(n, colptr, rowind, val);              /* :240-262 */
    int64_t nnz = colptr[n];
    free(dsum); free(cnt); free(pos);
    return nnz;
}

void orc_dependency_counts(int n, const int64_t *colptr, const int *rowind, int *dep)
{   /* pre_process.hpp:474-496: entries strictly above the diagonal */
    for (int k = 0; k < n; ++k) {
        int c = 0;
        for (int64_t p = colptr[k]; p < colptr[k + 1]; ++p)
            if (rowind[p] < k) ++c;
        dep[k] = c;
    }
}

/* ------------------------------------------------------------------------ */
/* Ground append: cpu_implementation/write_graph.jl:367-383                 */
/* ------------------------------------------------------------------------ */
int64_t orc_ground_append(int n, const int64_t *colptr, const int *rowind, const double *val,
                          int64_t *colptr2, int *rowind2, double *val2)
{
    double total = 0.0;
    double *app = (double *)malloc((size_t)n * sizeof(double));
    for (int j = 0; j < n; ++j) {
        double cs = 0.0;
        for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p) cs += val[p];
        total += cs;
        double a = -cs;
        app[j] = a > 0.0 ? 0.0 : a;                    /* :375-377 */
    }
    if (total < 0.0 && fabs(total) > 1e-4) { free(app); return -2; }   /* :369-372 */
    if (!(total > 1e-9)) { free(app); return 0; }                       /* :373 */
    int g = n;
    int64_t q = 0;
    double corner = 0.0;
    for (int j = 0; j < n; ++j) corner += app[j];
    corner = -corner;                                  /* :381 */
    colptr2[0] = 0;
    for (int j = 0; j < n; ++j) {
        for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p) { rowind2[q] = rowind[p]; val2[q] = val[p]; ++q; }
        if (app[j] != 0.0) { rowind2[q] = g; val2[q] = app[j]; ++q; }
        colptr2[j + 1] = q;
    }
    for (int j = 0; j < n; ++j)
        if (app[j] != 0.0) { rowind2[q] = j; val2[q] = app[j]; ++q; }
    rowind2[q] = g; val2[q] = corner; ++q;
    colptr2[n + 1] = q;
    free(app);
    return q;
}

int64_t orc_trim_last(int n, const int64_t *colptr, const int *rowind, const double *val,
                      int64_t *colptr2, int *rowind2, double *val2)
{   /* driver_physics.cu:1595-1734: drop row and column n-1 */
    int64_t q = 0;
    colptr2[0] = 0;
    for (int j = 0; j < n - 1; ++j) {
        for (int64_t p = colptr[j]; p < colptr[j + 1]; ++p)
            if (rowind[p] != n - 1) { rowind2[q] = rowind[p]; val2[q] = val[p]; ++q; }
        colptr2[j + 1] = q;
    }
    return q;
}

/* ------------------------------------------------------------------------ */
/* Orderings: write_graph.jl:46-78 (graph) and 333-365 (physics)            */
/* ------------------------------------------------------------------------ */
void orc_random_perm(int n, uint64_t seed, int *perm)
{
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int i = n - 1; i > 0; --i) {
        uint32_t o[4];
        philox_keyed(seed, (uint32_t)i, 0u, ORC_TAG_PERM, o);
        uint64_t r = ((uint64_t)o[1] << 32) | o[0];
        int j = (int)(((unsigned __int128)r * (unsigned __int128)(uint64_t)(i + 1)) >> 64);
        int t = perm[i]; perm[i] = perm[j]; perm[j] = t;
    }
}

void orc_nnz_sort_perm(int n, const int64_t *colptr, uint64_t seed, int *perm)
{
    int *p1 = (int *)malloc((size_t)n * sizeof(int));
    orc_random_perm(n, seed, p1);                      /* randperm, :47-48 */
    /* stable counting sort by column count of G[p1,p1] (sortperm, :56) */
    int64_t maxc = 0;
    for (int i = 0; i < n; ++i) {
        int64_t c = colptr[p1[i] + 1] - colptr[p1[i]];
        if (c > maxc) maxc = c;
    }
    int64_t *bucket = (int64_t *)calloc((size_t)maxc + 2, sizeof(int64_t));
    for (int i = 0; i < n; ++i) bucket[colptr[p1[i] + 1] - colptr[p1[i]] + 1]++;
    for (int64_t c = 0; c <= maxc; ++c) bucket[c + 1] += bucket[c];
    for (int i = 0; i < n; ++i) {
        int64_t c = colptr[p1[i] + 1] - colptr[p1[i]];
        perm[bucket[c]++] = p1[i];
    }
    free(bucket); free(p1);
}

int64_t orc_permute(int n, const int64_t *colptr, const int *rowind, const double *val,
                    const int *perm, int64_t *colptr2, int *rowind2, double *val2)
{
    int *inv = (int *)malloc((size_t)n * sizeof(int));
    for (int i = 0; i < n; ++i) inv[perm[i]] = i;
    int64_t q = 0;
    colptr2[0] = 0;
    for (int j = 0; j < n; ++j) {
        int oc = perm[j];
        for (int64_t p = colptr[oc]; p < colptr[oc + 1]; ++p) {
            rowind2[q] = inv[rowind[p]];
            val2[q] = val[p];
            ++q;
        }
        colptr2[j + 1] = q;
    }
    sort_columns(n, colptr2, rowind2, val2);
    free(inv);
    return q;
}

/* ------------------------------------------------------------------------ */
/* Randomized elimination: cpu_implementation/driver_local.cpp:302-707      */
/* ------------------------------------------------------------------------ */
typedef struct { int row; int next; double w; } fill_t;

/* Forward cumulative sums of e[0..d).w in the canonical blocked order shared
 * with the HIP kernel: blocks of 64 entries; inside a block the doubling scan
 * (offsets 1, 2, ..., 32, every step from the previous step's values: x_i +=
 * x_{i-off}); then the previous block's last sum is added.  The reference sums
 * sequentially on the CPU (driver_local.cpp:573-593) and with a CUB scan on
 * the GPU (driver.cu:931); any fixed order restates it, this one lets a wave
 * compute it in six shuffle steps and still match the oracle bit for bit. */
static void canonical_scan(const entry_t *e, int64_t d, double *C)
{
    double x[64], y[64];
    double carry = 0.0;
    for (int64_t b = 0; b < d; b += 64) {
        const int nb = (int)(d - b < 64 ? d - b : 64);
        for (int i = 0; i < 64; ++i) x[i] = i < nb ? e[b + i].w : 0.0;
        for (int off = 1; off < 64; off <<= 1) {
            for (int i = 0; i < 64; ++i) y[i] = i >= off ? x[i] + x[i - off] : x[i];
            memcpy(x, y, sizeof x);
        }
        for (int i = 0; i < nb; ++i) C[b + i] = b == 0 ? x[i] : carry + x[i];
        carry = C[b + nb - 1];
    }
}

int orc_factor(int n, const int64_t *colptr, const int *rowind, const double *val,
               uint64_t seed, int mutation, int64_t cap,
               int64_t *lt_ptr, int *lt_ind, double *lt_val, double *diag, int *level,
               orc_factor_stats *st)
{
    int *head = (int *)malloc((size_t)n * sizeof(int));
    int *nfill = (int *)calloc((size_t)n, sizeof(int));
    for (int i = 0; i < n; ++i) { head[i] = -1; level[i] = 0; }
    int64_t pool_cap = 1024, pool_n = 0;
    fill_t *pool = (fill_t *)malloc((size_t)pool_cap * sizeof(fill_t));
    int64_t work_cap = 64;
    entry_t *e = (entry_t *)malloc((size_t)work_cap * sizeof(entry_t));
    entry_t *byrow = (entry_t *)malloc((size_t)work_cap * sizeof(entry_t));
    double *C = (double *)malloc((size_t)work_cap * sizeof(double));
    orc_factor_stats s;
    memset(&s, 0, sizeof(s));
    int rc = 0;
    lt_ptr[0] = 0;

    for (int k = 0; k < n; ++k) {
        /* 1. gather: strict-lower original entries (as |v|) + pending fills
         *    (driver_local.cpp:459-487, cpu auxilliary.hpp:89-173) */
        int64_t norig = 0;
        for (int64_t p = colptr[k]; p < colptr[k + 1]; ++p)
            if (rowind[p] > k) ++norig;
        int64_t m = norig + nfill[k];
        if (m > work_cap) {
            while (work_cap < m) work_cap *= 2;
            e = (entry_t *)realloc(e, (size_t)work_cap * sizeof(entry_t));
            byrow = (entry_t *)realloc(byrow, (size_t)work_cap * sizeof(entry_t));
            C = (double *)realloc(C, (size_t)work_cap * sizeof(double));
        }
        int64_t t = 0;
        for (int64_t p = colptr[k]; p < colptr[k + 1]; ++p)
            if (rowind[p] > k) { e[t].row = rowind[p]; e[t].w = fabs(val[p]); e[t].mult = 1; ++t; }
        for (int f = head[k]; f >= 0; f = pool[f].next) {
            e[t].row = pool[f].row; e[t].w = pool[f].w; e[t].mult = 1; ++t;
        }
        s.gathered += m;
        if (m == 0) {                                   /* :491-498 */
            diag[k] = 0.0;
            lt_ptr[k + 1] = lt_ptr[k];
            s.zero_pivots++;
            continue;
        }
        /* 2. sort by (row, value) and merge duplicates left to right (:520-560) */
        qsort(e, (size_t)m, sizeof(entry_t), cmp_row_value);
        int64_t d = 0;
        for (int64_t i = 0; i < m; ++i) {
            if (d > 0 && e[i].row == e[d - 1].row) {
                if (mutation != ORC_MUT_NO_MERGE) e[d - 1].w += e[i].w;
                e[d - 1].mult += e[i].mult;
            } else {
                e[d++] = e[i];
            }
        }
        memcpy(byrow, e, (size_t)d * sizeof(entry_t));
        if (d > s.max_degree) s.max_degree = d;
        /* 3. sort by (value, row); forward cumulative sums (:573-593) */
        qsort(e, (size_t)d, sizeof(entry_t), cmp_value_row);
        canonical_scan(e, d, C);
        double S = C[d - 1];
        /* 4. one sampled fill per entry except the last (:598-631,
         *    gpu auxilliary.hpp:543-592) */
        for (int64_t i = 0; i + 1 < d; ++i) {
            double u = orc_uniform(seed, (uint32_t)i, (uint32_t)k, ORC_TAG_SAMPLE);
            double rest = S - C[i];
            double tgt = u * rest;
            int64_t lo = i + 1, hi = d - 1;                /* lower_bound on (C_j - C_i) >= tgt */
            while (lo < hi) {
                int64_t mid = lo + (hi - lo) / 2;
                if (C[mid] - C[i] < tgt) lo = mid + 1; else hi = mid;
            }
            int64_t j = lo;
            if (mutation == ORC_MUT_SAMPLE_SHIFT && j + 1 < d) j = j + 1;
            int rr = e[i].row > e[j].row ? e[i].row : e[j].row;
            int cc = e[i].row > e[j].row ? e[j].row : e[i].row;
            double fw = (mutation == ORC_MUT_FILL_WEIGHT) ? e[i].w : (e[i].w * rest) / S;
            if (pool_n == pool_cap) {
                pool_cap *= 2;
                pool = (fill_t *)realloc(pool, (size_t)pool_cap * sizeof(fill_t));
            }
            pool[pool_n].row = rr; pool[pool_n].w = fw; pool[pool_n].next = head[cc];
            head[cc] = (int)pool_n;
            ++pool_n;
            nfill[cc]++;
            s.fills++;
        }
        /* 5. factor column, row order, values -w/S; pivot S (:638-645, 881-934) */
        if (lt_ptr[k] + d > cap) { rc = -1; break; }
        for (int64_t i = 0; i < d; ++i) {
            lt_ind[lt_ptr[k] + i] = byrow[i].row;
            lt_val[lt_ptr[k] + i] = -(byrow[i].w / S);
            int lv = level[k] + 1;
            if (level[byrow[i].row] < lv) level[byrow[i].row] = lv;
        }
        lt_ptr[k + 1] = lt_ptr[k] + d;
        diag[k] = (mutation == ORC_MUT_PIVOT && d > 1) ? C[d - 2] : S;
        s.nnz_off += d;
    }
    if (rc == 0) {
        int mx = -1;
        for (int i = 0; i < n; ++i) if (level[i] > mx) mx = level[i];
        s.nlevels = mx + 1;
    }
    if (st) *st = s;
    free(head); free(nfill); free(pool); free(e); free(byrow); free(C);
    return rc;
}

int64_t orc_factor_trim_last(int n, const int64_t *lt_ptr, const int *lt_ind, const double *lt_val,
                             int64_t *lt_ptr2, int *lt_ind2, double *lt_val2)
{   /* driver_physics.cu:1945-2010 (remove_last_csr + shift_element) */
    int64_t q = 0;
    lt_ptr2[0] = 0;
    for (int k = 0; k < n - 1; ++k) {
        for (int64_t p = lt_ptr[k]; p < lt_ptr[k + 1]; ++p)
            if (lt_ind[p] != n - 1) { lt_ind2[q] = lt_ind[p]; lt_val2[q] = lt_val[p]; ++q; }
        lt_ptr2[k + 1] = q;
    }
    return q;
}

/* ------------------------------------------------------------------------ */
/* SpMV, preconditioner, PCG: gpu_implementation/solver.hpp                  */
/* ------------------------------------------------------------------------ */
void orc_spmv(int n, const int64_t *ptr, const int *ind, const double *val, const double *x, double *y)
{
    for (int i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int64_t p = ptr[i]; p < ptr[i + 1]; ++p) acc = fma(val[p], x[ind[p]], acc);
        y[i] = acc;
    }
}

void orc_precond(int n, const int64_t *lt_ptr, const int *lt_ind, const double *lt_val,
                 const double *diag, int precision, int mutation, const double *r, double *z)
{   /* solver.hpp:149-233: SpSV(L) -> diagonal_solve -> SpSV(L^T) */
    if (precision == 32) {
        float *y = (float *)malloc((size_t)n * sizeof(float));
        float *zz = (float *)malloc((size_t)n * sizeof(float));
        for (int i = 0; i < n; ++i) y[i] = (float)r[i];
        for (int k = 0; k < n; ++k) {
            float yk = y[k];
            for (int64_t p = lt_ptr[k]; p < lt_ptr[k + 1]; ++p)
                y[lt_ind[p]] = fmaf(-(float)lt_val[p], yk, y[lt_ind[p]]);
        }
        for (int k = 0; k < n; ++k) { float dk = (float)diag[k]; if (dk != 0.0f) y[k] = y[k] / dk; }
        for (int k = n - 1; k >= 0; --k) {
            float acc = y[k];
            for (int64_t p = lt_ptr[k]; p < lt_ptr[k + 1]; ++p) {
                float l = (float)lt_val[p];
                acc = (mutation == ORC_MUT_PRECOND_SIGN) ? fmaf(l, zz[lt_ind[p]], acc)
                                                         : fmaf(-l, zz[lt_ind[p]], acc);
            }
            zz[k] = acc;
        }
        for (int i = 0; i < n; ++i) z[i] = (double)zz[i];
        free(y); free(zz);
        return;
    }
    double *y = (double *)malloc((size_t)n * sizeof(double));
    memcpy(y, r, (size_t)n * sizeof(double));
    for (int k = 0; k < n; ++k) {
        double yk = y[k];
        for (int64_t p = lt_ptr[k]; p < lt_ptr[k + 1]; ++p)
            y[lt_ind[p]] = fma(-lt_val[p], yk, y[lt_ind[p]]);
    }
    for (int k = 0; k < n; ++k) if (diag[k] != 0.0) y[k] = y[k] / diag[k];   /* :64-84 */
    for (int k = n - 1; k >= 0; --k) {
        double acc = y[k];
        for (int64_t p = lt_ptr[k]; p < lt_ptr[k + 1]; ++p)
            acc = (mutation == ORC_MUT_PRECOND_SIGN) ? fma(lt_val[p], z[lt_ind[p]], acc)
                                                     : fma(-lt_val[p], z[lt_ind[p]], acc);
        z[k] = acc;
    }
    free(y);
}

static double dot(int n, const double *a, const double *b)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

int orc_pcg(int n, const int64_t *a_ptr, const int *a_ind, const double *a_val,
            const int64_t *lt_ptr, const int *lt_ind, const double *lt_val, const double *diag,
            int precision, const double *b, double *x, double tol, int maxit, int stop_rule,
            orc_pcg_stats *st)
{
    double *r = (double *)malloc((size_t)n * sizeof(double));
    double *z = (double *)malloc((size_t)n * sizeof(double));
    double *p = (double *)malloc((size_t)n * sizeof(double));
    double *ap = (double *)malloc((size_t)n * sizeof(double));
    int rc = 0;
    for (int i = 0; i < n; ++i) { x[i] = 0.0; r[i] = b[i]; }          /* :268-277, 462-466 */
    orc_precond(n, lt_ptr, lt_ind, lt_val, diag, precision, 0, r, z);   /* :470-473 */
    memcpy(p, z, (size_t)n * sizeof(double));
    double r1 = dot(n, r, z), r0;
    double bnorm = sqrt(dot(n, b, b));
    double r1_0 = r1;
    int k = 0;
    int go = (stop_rule == ORC_STOP_REFERENCE) ? (r1 > tol)             /* :493 */
                                               : (bnorm > 0.0 && sqrt(dot(n, r, r)) / bnorm > tol);
    while (k < maxit && go) {
        orc_spmv(n, a_ptr, a_ind, a_val, p, ap);
        double pap = dot(n, p, ap);
        if (!(pap > 0.0) || !isfinite(pap)) { rc = -3; break; }
        double alpha = r1 / pap;
        for (int i = 0; i < n; ++i) { x[i] = fma(alpha, p[i], x[i]); r[i] = fma(-alpha, ap[i], r[i]); }
        orc_precond(n, lt_ptr, lt_ind, lt_val, diag, precision, 0, r, z);
        r0 = r1;
        r1 = dot(n, r, z);
        double beta = r1 / r0;
        for (int i = 0; i < n; ++i) { double t = beta * p[i]; p[i] = t + z[i]; }   /* :530-531 */
        ++k;
        go = (stop_rule == ORC_STOP_REFERENCE) ? (sqrt(r1) / sqrt(r1_0) > tol)    /* :536 */
                                               : (sqrt(dot(n, r, r)) / bnorm > tol);
    }
    if (st) {
        st->iterations = k;
        st->converged = (rc == 0 && !go);
        st->rz0 = r1_0;
        st->rz = r1;
        st->rel_res_rec = bnorm > 0.0 ? sqrt(dot(n, r, r)) / bnorm : 0.0;
        orc_spmv(n, a_ptr, a_ind, a_val, x, ap);                          /* :549-569 */
        for (int i = 0; i < n; ++i) ap[i] -= b[i];
        st->rel_res_true = bnorm > 0.0 ? sqrt(dot(n, ap, ap)) / bnorm : 0.0;
    }
    free(r); free(z); free(p); free(ap);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* IC(0): the context preconditioner (see rchol_oracle.h)                    */
/* ------------------------------------------------------------------------ */
int orc_ic0(int n, const int64_t *colptr, const int *rowind, const double *val, int mutation, int64_t cap,
            int64_t *lt_ptr, int *lt_ind, double *lt_val, double *diag, int *level)
{
    /* row i of L = rows j < i of column i (the pattern is symmetric) */
    int64_t *up = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    for (int i = 0; i < n; ++i) {
        int64_t c = 0;
        for (int64_t p = colptr[i]; p < colptr[i + 1]; ++p) c += rowind[p] < i;
        up[i + 1] = up[i] + c;
    }
    int *ui = (int *)malloc((size_t)(up[n] > 0 ? up[n] : 1) * sizeof(int));
    double *uv = (double *)malloc((size_t)(up[n] > 0 ? up[n] : 1) * sizeof(double));
    double *l = (double *)malloc((size_t)(up[n] > 0 ? up[n] : 1) * sizeof(double));
    for (int i = 0; i < n; ++i) {
        int64_t q = up[i];
        for (int64_t p = colptr[i]; p < colptr[i + 1]; ++p)
            if (rowind[p] < i) { ui[q] = rowind[p]; uv[q] = val[p]; ++q; }
    }
    for (int i = 0; i < n; ++i) {
        double aii = 0.0;
        for (int64_t p = colptr[i]; p < colptr[i + 1]; ++p)
            if (rowind[p] == i) aii += val[p];
        double di = aii;
        int lev = 0;
        for (int64_t t = up[i]; t < up[i + 1]; ++t) {
            const int j = ui[t];
            double s = uv[t];
            /* common neighbours m < j in ascending order: each m of row j is
             * looked up in row i's entries before t by bisection (a dense row,
             * such as an appended ground vertex, would make a linear merge
             * quadratic in its length) */
            int64_t p = up[i];
            for (int64_t q = up[j]; q < up[j + 1]; ++q) {
                const int mq = ui[q];
                int64_t hi = t;
                while (p < hi) {
                    const int64_t mid = p + ((hi - p) >> 1);
                    if (ui[mid] < mq) p = mid + 1; else hi = mid;
                }
                if (p < t && ui[p] == mq) {
                    s -= mutation == ORC_MUT_IC0_NO_PIVOT ? l[p] * l[q] : l[p] * l[q] * diag[mq];
                    ++p;
                }
            }
            const double dj = diag[j];
            const double lij = dj > 0.0 ? s / dj : 0.0;
            l[t] = lij;
            di -= lij * lij * dj;
            if (level[j] + 1 > lev) lev = level[j] + 1;
        }
        if (!(di > 1e-14 * aii)) di = 0.0;
        diag[i] = di;
        level[i] = lev;
    }
    /* columns of L (rows > k, ascending) */
    lt_ptr[0] = 0;
    for (int k = 0; k < n; ++k) {
        int64_t c = 0;
        for (int64_t p = colptr[k]; p < colptr[k + 1]; ++p) c += rowind[p] > k;
        lt_ptr[k + 1] = lt_ptr[k] + c;
    }
    int rc = 0;
    if (lt_ptr[n] > cap) rc = -1;
    if (rc == 0) {
        int64_t *pos = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
        memcpy(pos, lt_ptr, (size_t)n * sizeof(int64_t));
        /* rows i ascending, so every column receives its rows in order */
        for (int i = 0; i < n; ++i)
            for (int64_t t = up[i]; t < up[i + 1]; ++t) {
                const int64_t q = pos[ui[t]]++;
                lt_ind[q] = i;
                lt_val[q] = l[t];
            }
        free(pos);
    }
    free(up); free(ui); free(uv); free(l);
    return rc;
}
