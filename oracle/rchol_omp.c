/*
 * rchol_omp.c -- multi-threaded CPU restatement of the same path (test
 * infrastructure and CPU baseline only; see rchol_oracle.h).
 *
 * orc_factor_omp: the reference CPU factorization is an OpenMP team that
 * eliminates columns as their dependencies retire
 * (cpu_implementation/driver_local.cpp:302-707, 802-829).  This restatement
 * does the same: a column is ready once every lower-numbered neighbour it
 * still has has been eliminated; threads take ready columns from a shared
 * queue, gather the column's original lower entries and pending fills, merge,
 * sort, sample and push the new fills onto the target columns' lists with an
 * atomic exchange of the list head.  The per-column steps are the oracle's
 * (same canonical merge and sampling orders, same Philox draws), so the factor
 * is bitwise equal to orc_factor for any thread count and any schedule.
 *
 * orc_pcg_omp: the reference PCG (gpu_implementation/solver.hpp:457-569) with
 * OpenMP SpMV, dots and axpys, and level-scheduled triangular solves (rows of
 * one level in parallel, levels in order: the CPU analogue of the SpSV
 * analysis the reference's cuSPARSE path does, solver.hpp:317-371).  Dots are
 * reduced in a fixed blocked order, so results do not depend on the thread
 * count; they differ from orc_pcg's sequential sums by rounding only.
 */
#include "rchol_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

typedef struct { int row; int next; double w; } ofill_t;
typedef struct { int row; int mult; double w; } oentry_t;

static int ocmp_row_value(const void *pa, const void *pb)
{
    const oentry_t *a = (const oentry_t *)pa, *b = (const oentry_t *)pb;
    if (a->row != b->row) return a->row < b->row ? -1 : 1;
    if (a->w != b->w) return a->w < b->w ? -1 : 1;
    return 0;
}

static int ocmp_value_row(const void *pa, const void *pb)
{
    const oentry_t *a = (const oentry_t *)pa, *b = (const oentry_t *)pb;
    if (a->w != b->w) return a->w < b->w ? -1 : 1;
    if (a->row != b->row) return a->row < b->row ? -1 : 1;
    return 0;
}

/* the canonical blocked forward sums of orc_factor */
static void oscan(const oentry_t *e, int64_t d, double *C)
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

enum { OMP_POOL_CHUNK = 4096 };

int orc_factor_omp(int n, const int64_t *colptr, const int *rowind, const double *val, uint64_t seed,
                   int threads, int64_t cap, int64_t *lt_ptr, int *lt_ind, double *lt_val, double *diag,
                   int *level, orc_factor_stats *st)
{
    if (threads <= 0) threads = omp_get_max_threads();
    int *head = (int *)malloc((size_t)n * sizeof(int));
    int *dep = (int *)calloc((size_t)n, sizeof(int));
    int *ready = (int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    int64_t *ostart = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    int *ocnt = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
    int64_t nnz_low = 0;
    for (int k = 0; k < n; ++k) {
        head[k] = -1;
        level[k] = 0;
        ready[k] = -1;
        for (int64_t p = colptr[k]; p < colptr[k + 1]; ++p)
            if (rowind[p] > k) { dep[rowind[p]]++; ++nnz_low; }
    }
    /* fill records: every column k emits d_k - 1 fills and d_k <= its
     * gathered count, so the pool never needs more than the factor's entries */
    int64_t pool_cap = cap + (int64_t)threads * OMP_POOL_CHUNK + 1024;
    ofill_t *pool = (ofill_t *)malloc((size_t)pool_cap * sizeof(ofill_t));
    int *outrow = (int *)malloc((size_t)(cap > 0 ? cap : 1) * sizeof(int));
    double *outval = (double *)malloc((size_t)(cap > 0 ? cap : 1) * sizeof(double));
    int64_t pool_top = 0, out_top = 0;
    int q_tail = 0, q_head = 0, done = 0, fail = 0;
    for (int k = 0; k < n; ++k)
        if (dep[k] == 0) ready[q_tail++] = k;
    orc_factor_stats tot;
    memset(&tot, 0, sizeof(tot));

#pragma omp parallel num_threads(threads)
    {
        int64_t wcap = 64;
        oentry_t *e = (oentry_t *)malloc((size_t)wcap * sizeof(oentry_t));
        oentry_t *byrow = (oentry_t *)malloc((size_t)wcap * sizeof(oentry_t));
        double *C = (double *)malloc((size_t)wcap * sizeof(double));
        int64_t pcur = 0, pend = 0;
        orc_factor_stats s;
        memset(&s, 0, sizeof(s));
        while (1) {
            if (__atomic_load_n(&fail, __ATOMIC_ACQUIRE)) break;
            if (__atomic_load_n(&done, __ATOMIC_ACQUIRE) >= n) break;
            int slot = __atomic_load_n(&q_head, __ATOMIC_RELAXED);
            if (slot >= n) break;
            if (slot >= __atomic_load_n(&q_tail, __ATOMIC_ACQUIRE)) continue;     /* nothing ready yet */
            if (!__atomic_compare_exchange_n(&q_head, &slot, slot + 1, 0, __ATOMIC_ACQ_REL, __ATOMIC_RELAXED))
                continue;
            int k;
            while ((k = __atomic_load_n(&ready[slot], __ATOMIC_ACQUIRE)) < 0) {}   /* slot being written */

            /* 1. gather */
            int64_t m = 0;
            for (int64_t p = colptr[k]; p < colptr[k + 1]; ++p)
                if (rowind[p] > k) ++m;
            for (int f = __atomic_load_n(&head[k], __ATOMIC_ACQUIRE); f >= 0; f = pool[f].next) ++m;
            if (m > wcap) {
                while (wcap < m) wcap *= 2;
                e = (oentry_t *)realloc(e, (size_t)wcap * sizeof(oentry_t));
                byrow = (oentry_t *)realloc(byrow, (size_t)wcap * sizeof(oentry_t));
                C = (double *)realloc(C, (size_t)wcap * sizeof(double));
            }
            int64_t t = 0;
            for (int64_t p = colptr[k]; p < colptr[k + 1]; ++p)
                if (rowind[p] > k) { e[t].row = rowind[p]; e[t].w = fabs(val[p]); e[t].mult = 1; ++t; }
            for (int f = __atomic_load_n(&head[k], __ATOMIC_ACQUIRE); f >= 0; f = pool[f].next) {
                e[t].row = pool[f].row; e[t].w = pool[f].w; e[t].mult = 1; ++t;
            }
            s.gathered += m;
            if (m == 0) {
                diag[k] = 0.0;
                ocnt[k] = 0;
                ostart[k] = 0;
                s.zero_pivots++;
                __atomic_fetch_add(&done, 1, __ATOMIC_ACQ_REL);
                continue;
            }
            /* 2. merge duplicate rows */
            qsort(e, (size_t)m, sizeof(oentry_t), ocmp_row_value);
            int64_t d = 0;
            for (int64_t i = 0; i < m; ++i) {
                if (d > 0 && e[i].row == e[d - 1].row) {
                    e[d - 1].w += e[i].w;
                    e[d - 1].mult += e[i].mult;
                } else {
                    e[d++] = e[i];
                }
            }
            memcpy(byrow, e, (size_t)d * sizeof(oentry_t));
            if (d > s.max_degree) s.max_degree = d;
            /* 3. sort by value, prefix sums */
            qsort(e, (size_t)d, sizeof(oentry_t), ocmp_value_row);
            oscan(e, d, C);
            const double S = C[d - 1];
            /* factor column (rows ascending) */
            const int64_t ob = __atomic_fetch_add(&out_top, d, __ATOMIC_RELAXED);
            if (ob + d > cap) { __atomic_store_n(&fail, 1, __ATOMIC_RELEASE); break; }
            for (int64_t i = 0; i < d; ++i) {
                outrow[ob + i] = byrow[i].row;
                outval[ob + i] = -(byrow[i].w / S);
            }
            ostart[k] = ob;
            ocnt[k] = (int)d;
            diag[k] = S;
            s.nnz_off += d;
            const int lk = level[k] + 1;
            for (int64_t i = 0; i < d; ++i) {       /* forward-solve level of the neighbours */
                int *lp = &level[byrow[i].row];
                int cur = __atomic_load_n(lp, __ATOMIC_RELAXED);
                while (cur < lk && !__atomic_compare_exchange_n(lp, &cur, lk, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {}
            }
            /* 4. sampled fills: (max, min) goes to column min, adds a pending
             *    lower neighbour to max */
            for (int64_t i = 0; i + 1 < d; ++i) {
                const double u = orc_uniform(seed, (uint32_t)i, (uint32_t)k, ORC_TAG_SAMPLE);
                const double rest = S - C[i];
                const double tgt = u * rest;
                int64_t lo = i + 1, hi = d - 1;
                while (lo < hi) {
                    const int64_t mid = lo + (hi - lo) / 2;
                    if (C[mid] - C[i] < tgt) lo = mid + 1; else hi = mid;
                }
                const int rr = e[i].row > e[lo].row ? e[i].row : e[lo].row;
                const int cc = e[i].row > e[lo].row ? e[lo].row : e[i].row;
                if (pcur == pend) {
                    pcur = __atomic_fetch_add(&pool_top, (int64_t)OMP_POOL_CHUNK, __ATOMIC_RELAXED);
                    pend = pcur + OMP_POOL_CHUNK;
                    if (pend > pool_cap) { __atomic_store_n(&fail, 1, __ATOMIC_RELEASE); break; }
                }
                const int64_t f = pcur++;
                pool[f].row = rr;
                pool[f].w = (e[i].w * rest) / S;
                __atomic_fetch_add(&dep[rr], 1, __ATOMIC_RELAXED);
                int old = __atomic_load_n(&head[cc], __ATOMIC_RELAXED);
                do {
                    pool[f].next = old;
                } while (!__atomic_compare_exchange_n(&head[cc], &old, (int)f, 1, __ATOMIC_RELEASE, __ATOMIC_RELAXED));
                s.fills++;
            }
            if (__atomic_load_n(&fail, __ATOMIC_ACQUIRE)) break;
            /* 5. retire: every neighbour loses this column as a lower
             *    neighbour (with its merged multiplicity) */
            for (int64_t i = 0; i < d; ++i) {
                const int r = byrow[i].row;
                if (__atomic_sub_fetch(&dep[r], byrow[i].mult, __ATOMIC_ACQ_REL) == 0) {
                    const int qs = __atomic_fetch_add(&q_tail, 1, __ATOMIC_ACQ_REL);
                    __atomic_store_n(&ready[qs], r, __ATOMIC_RELEASE);
                }
            }
            __atomic_fetch_add(&done, 1, __ATOMIC_ACQ_REL);
        }
#pragma omp critical
        {
            tot.nnz_off += s.nnz_off;
            tot.fills += s.fills;
            tot.gathered += s.gathered;
            if (s.max_degree > tot.max_degree) tot.max_degree = s.max_degree;
            tot.zero_pivots += s.zero_pivots;
        }
        free(e); free(byrow); free(C);
    }
    int rc = fail ? -1 : 0;
    if (rc == 0) {
        /* CSR of L^T in column order */
        lt_ptr[0] = 0;
        for (int k = 0; k < n; ++k) lt_ptr[k + 1] = lt_ptr[k] + ocnt[k];
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int k = 0; k < n; ++k) {
            memcpy(lt_ind + lt_ptr[k], outrow + ostart[k], (size_t)ocnt[k] * sizeof(int));
            memcpy(lt_val + lt_ptr[k], outval + ostart[k], (size_t)ocnt[k] * sizeof(double));
        }
        int mx = -1;
        for (int i = 0; i < n; ++i) if (level[i] > mx) mx = level[i];
        tot.nlevels = mx + 1;
    }
    (void)nnz_low;
    if (st) *st = tot;
    free(head); free(dep); free(ready); free(ostart); free(ocnt); free(pool); free(outrow); free(outval);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* PCG                                                                      */
/* ------------------------------------------------------------------------ */
enum { OMP_DOT_BLOCK = 8192 };

/* dot in a fixed order: blocks of OMP_DOT_BLOCK summed sequentially, block
 * sums added in block order */
static double odot(int n, const double *a, const double *b, double *part, int threads)
{
    const int nb = (n + OMP_DOT_BLOCK - 1) / OMP_DOT_BLOCK;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int q = 0; q < nb; ++q) {
        const int lo = q * OMP_DOT_BLOCK, hi = lo + OMP_DOT_BLOCK < n ? lo + OMP_DOT_BLOCK : n;
        double s = 0.0;
        for (int i = lo; i < hi; ++i) s += a[i] * b[i];
        part[q] = s;
    }
    double s = 0.0;
    for (int q = 0; q < nb; ++q) s += part[q];
    return s;
}

static void ospmv(int n, const int64_t *ptr, const int *ind, const double *val, const double *x, double *y,
                  int threads)
{
#pragma omp parallel for num_threads(threads) schedule(static, 1024)
    for (int i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int64_t p = ptr[i]; p < ptr[i + 1]; ++p) acc = fma(val[p], x[ind[p]], acc);
        y[i] = acc;
    }
}

typedef struct {
    int n, nlev;
    int *lev_ptr, *lev_rows;    /* rows grouped by forward level */
    int64_t *l_ptr;             /* CSR of L (row i: columns j < i), from L^T */
    int *l_ind;
    double *l_val;
} otri_t;

static void otri_build(int n, const int64_t *lt_ptr, const int *lt_ind, const double *lt_val, const int *level,
                       otri_t *T, int threads)
{
    T->n = n;
    int mx = -1;
    for (int i = 0; i < n; ++i) if (level[i] > mx) mx = level[i];
    T->nlev = mx + 1;
    T->lev_ptr = (int *)calloc((size_t)T->nlev + 1, sizeof(int));
    T->lev_rows = (int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    for (int i = 0; i < n; ++i) T->lev_ptr[level[i] + 1]++;
    for (int l = 0; l < T->nlev; ++l) T->lev_ptr[l + 1] += T->lev_ptr[l];
    int *cur = (int *)malloc((size_t)(T->nlev > 0 ? T->nlev : 1) * sizeof(int));
    memcpy(cur, T->lev_ptr, (size_t)T->nlev * sizeof(int));
    for (int i = 0; i < n; ++i) T->lev_rows[cur[level[i]]++] = i;
    free(cur);
    /* transpose L^T (rows k: entries r > k) into rows of L, columns ascending */
    const int64_t nnz = lt_ptr[n];
    T->l_ptr = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    T->l_ind = (int *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int));
    T->l_val = (double *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
    for (int64_t p = 0; p < nnz; ++p) T->l_ptr[lt_ind[p] + 1]++;
    for (int i = 0; i < n; ++i) T->l_ptr[i + 1] += T->l_ptr[i];
    int64_t *pos = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    memcpy(pos, T->l_ptr, (size_t)n * sizeof(int64_t));
    for (int k = 0; k < n; ++k)
        for (int64_t p = lt_ptr[k]; p < lt_ptr[k + 1]; ++p) {
            const int64_t q = pos[lt_ind[p]]++;
            T->l_ind[q] = k;
            T->l_val[q] = lt_val[p];
        }
    free(pos);
    (void)threads;
}

static void otri_free(otri_t *T)
{
    free(T->lev_ptr); free(T->lev_rows); free(T->l_ptr); free(T->l_ind); free(T->l_val);
}

/* z = L^-T D+ L^-1 r, level by level (forward: levels ascending, pull over
 * rows of L; backward: levels descending, pull over rows of L^T) */
static void oprecond(const otri_t *T, const int64_t *lt_ptr, const int *lt_ind, const double *lt_val,
                     const double *diag, const double *r, double *y, double *z, int threads)
{
#pragma omp parallel num_threads(threads)
    {
        for (int l = 0; l < T->nlev; ++l) {
#pragma omp for schedule(static)
            for (int q = T->lev_ptr[l]; q < T->lev_ptr[l + 1]; ++q) {
                const int i = T->lev_rows[q];
                double acc = r[i];
                for (int64_t p = T->l_ptr[i]; p < T->l_ptr[i + 1]; ++p) acc = fma(-T->l_val[p], y[T->l_ind[p]], acc);
                y[i] = acc;
            }
        }
        for (int l = T->nlev - 1; l >= 0; --l) {
#pragma omp for schedule(static)
            for (int q = T->lev_ptr[l]; q < T->lev_ptr[l + 1]; ++q) {
                const int i = T->lev_rows[q];
                double acc = diag[i] != 0.0 ? y[i] / diag[i] : y[i];
                for (int64_t p = lt_ptr[i]; p < lt_ptr[i + 1]; ++p) acc = fma(-lt_val[p], z[lt_ind[p]], acc);
                z[i] = acc;
            }
        }
    }
}

int orc_pcg_omp(int n, const int64_t *a_ptr, const int *a_ind, const double *a_val, const int64_t *lt_ptr,
                const int *lt_ind, const double *lt_val, const double *diag, const int *level, const double *b,
                double *x, double tol, int maxit, int threads, double *setup_s, orc_pcg_stats *st)
{
    if (threads <= 0) threads = omp_get_max_threads();
    const double t0 = omp_get_wtime();
    otri_t T;
    otri_build(n, lt_ptr, lt_ind, lt_val, level, &T, threads);
    double *r = (double *)malloc((size_t)n * sizeof(double));
    double *z = (double *)malloc((size_t)n * sizeof(double));
    double *y = (double *)malloc((size_t)n * sizeof(double));
    double *p = (double *)malloc((size_t)n * sizeof(double));
    double *ap = (double *)malloc((size_t)n * sizeof(double));
    double *part = (double *)malloc((size_t)(n / OMP_DOT_BLOCK + 2) * sizeof(double));
    if (setup_s) *setup_s = omp_get_wtime() - t0;
    int rc = 0;
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int i = 0; i < n; ++i) { x[i] = 0.0; r[i] = b[i]; }
    oprecond(&T, lt_ptr, lt_ind, lt_val, diag, r, y, z, threads);
    memcpy(p, z, (size_t)n * sizeof(double));
    double r1 = odot(n, r, z, part, threads), r0;
    const double bnorm = sqrt(odot(n, b, b, part, threads));
    const double r1_0 = r1;
    int k = 0;
    int go = r1 > tol;                                                  /* solver.hpp:493 */
    while (k < maxit && go) {
        ospmv(n, a_ptr, a_ind, a_val, p, ap, threads);
        const double pap = odot(n, p, ap, part, threads);
        if (!(pap > 0.0) || !isfinite(pap)) { rc = -3; break; }
        const double alpha = r1 / pap;
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int i = 0; i < n; ++i) { x[i] = fma(alpha, p[i], x[i]); r[i] = fma(-alpha, ap[i], r[i]); }
        oprecond(&T, lt_ptr, lt_ind, lt_val, diag, r, y, z, threads);
        r0 = r1;
        r1 = odot(n, r, z, part, threads);
        const double beta = r1 / r0;
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int i = 0; i < n; ++i) { const double t = beta * p[i]; p[i] = t + z[i]; }
        ++k;
        go = sqrt(r1) / sqrt(r1_0) > tol;                                /* solver.hpp:536 */
    }
    if (st) {
        st->iterations = k;
        st->converged = (rc == 0 && !go);
        st->rz0 = r1_0;
        st->rz = r1;
        st->rel_res_rec = bnorm > 0.0 ? sqrt(odot(n, r, r, part, threads)) / bnorm : 0.0;
        ospmv(n, a_ptr, a_ind, a_val, x, ap, threads);                   /* solver.hpp:549-569 */
        for (int i = 0; i < n; ++i) ap[i] -= b[i];
        st->rel_res_true = bnorm > 0.0 ? sqrt(odot(n, ap, ap, part, threads)) / bnorm : 0.0;
    }
    otri_free(&T);
    free(r); free(z); free(y); free(p); free(ap); free(part);
    return rc;
}
