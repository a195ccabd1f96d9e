/*
 * rchol_oracle.h -- CPU oracle for the randomized-Cholesky factor + PCG path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, single-threaded restatement of
 * the reference algorithm used to check the HIP implementation.  Only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark may load it,
 * and only as a checker; the product library never links or calls it.
 *
 * Reference files followed (paths relative to the reference repository):
 *   cpu_implementation/driver_local.cpp:302-707   randomized elimination
 *   cpu_implementation/driver_local.cpp:881-956   factor assembly (CSR of L^T)
 *   gpu_implementation/pre_process.hpp:189-306    Laplacian construction
 *   gpu_implementation/pre_process.hpp:474-496    dependency counts
 *   gpu_implementation/solver.hpp:63-84,149-233   preconditioner apply
 *   gpu_implementation/solver.hpp:236-593         PCG recurrence / stop rules
 *   gpu_implementation/driver_physics.cu:1595-1734,1945-2090  ground trimming
 *   cpu_implementation/write_graph.jl:43-95,328-414  orderings, ground append
 *
 * Canonicalisation (documented deviations that make the factor a pure
 * function of (matrix, seed) and hence bit-comparable with the GPU kernel):
 *   - merge order: gathered entries sorted by (row, value), duplicate rows
 *     summed left to right;
 *   - sampling order: merged entries sorted by (value, row);
 *   - random draws: Philox4x32-10 keyed by the seed, counter (i, column, tag);
 *     u in (0, 1].
 *
 * Parity status of each function is stated in its comment ("pinned by ...").
 */
#ifndef RCHOL_ORACLE_H
#define RCHOL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x32-10 block function (Salmon, Moraes, Dror, Shaw, SC'11).
 * Pinned by the Random123 known-answer vectors in the tests. */
void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* u in (0,1] from one Philox block keyed by seed, counter (a, b, tag, 0).
 * 53-bit mantissa: ((hi:lo) >> 11) + 1, times 2^-53. */
double orc_uniform(uint64_t seed, uint32_t a, uint32_t b, uint32_t tag);

/* Counter tags: independent Philox streams for the different consumers. */
enum { ORC_TAG_SAMPLE = 0, ORC_TAG_NORMAL = 1, ORC_TAG_PERM = 2, ORC_TAG_WEIGHT = 3, ORC_TAG_POINT = 4 };

/* Standard normal by Box-Muller (cosine branch) from one Philox block. */
double orc_normal(uint64_t seed, uint32_t index, uint32_t stream);

/* out[i] = orc_normal(seed, i, stream) for i < n. */
void orc_normal_vector(uint64_t seed, int64_t n, uint32_t stream, double *out);

/* Laplacian construction from MM triplets (pre_process.hpp:189-306, 322-357):
 * off-diagonals become -|v|, diagonals +|v|; if the input holds no diagonal
 * entry at all, a diagonal equal to the column sum of |off-diagonals| is
 * appended.  Entries are sorted by (col, row).  Explicit zero off-diagonal
 * entries are dropped (the reference warns that they break the algorithm,
 * cpu pre_process.hpp:816-820).  Output arrays need capacity m + n.
 * Returns the nnz written, or -1 on an index out of range.
 * Pinned by closed forms on hand-made graphs (tests/golden). */
int64_t orc_laplacian_from_triplets(int n, int64_t m, const int *rows, const int *cols,
                                    const double *vals, int64_t *colptr, int *rowind,
                                    double *val);

/* Dependency counts dep[k] = #{u < k : A[u,k] != 0} (pre_process.hpp:474-496). */
void orc_dependency_counts(int n, const int64_t *colptr, const int *rowind, int *dep);

/* SDDM -> Laplacian by an appended ground vertex (write_graph.jl:367-383).
 * Output has n+1 vertices, ground last.  Capacity: nnz + 2n + 1.
 * Returns nnz written; 0 if the matrix already sums to zero (no append,
 * output untouched); -2 if not diagonally dominant (sum < -1e-4). */
int64_t orc_ground_append(int n, const int64_t *colptr, const int *rowind, const double *val,
                          int64_t *colptr2, int *rowind2, double *val2);

/* Remove the last row and column (driver_physics.cu:1595-1734). Returns nnz. */
int64_t orc_trim_last(int n, const int64_t *colptr, const int *rowind, const double *val,
                      int64_t *colptr2, int *rowind2, double *val2);

/* Orderings (write_graph.jl:43-95 / 328-365).  perm[new] = old.
 * random: Fisher-Yates driven by Philox(seed).
 * nnz-sort: random pre-permutation, then a stable sort by column count.
 * Pinned by invariants only (the reference's Julia randperm is unseeded). */
void orc_random_perm(int n, uint64_t seed, int *perm);
void orc_nnz_sort_perm(int n, const int64_t *colptr, uint64_t seed, int *perm);

/* B = A[perm, perm]; columns of B sorted by row.  Returns nnz. */
int64_t orc_permute(int n, const int64_t *colptr, const int *rowind, const double *val,
                    const int *perm, int64_t *colptr2, int *rowind2, double *val2);

/* Planted mistakes for tests that show the pins have teeth (0 = none). */
enum {
    ORC_MUT_NONE = 0,
    ORC_MUT_FILL_WEIGHT = 1,   /* fill weight w_i instead of w_i (S-C_i)/S      */
    ORC_MUT_SAMPLE_SHIFT = 2,  /* sampled index j+1 (clamped) instead of j       */
    ORC_MUT_NO_MERGE = 3,      /* keep only the first entry of each duplicate row */
    ORC_MUT_PIVOT = 4,         /* D_k = S without the last entry                 */
    ORC_MUT_PRECOND_SIGN = 5,  /* backward solve adds instead of subtracting     */
    ORC_MUT_IC0_NO_PIVOT = 6   /* IC(0): d_m left out of the update sums          */
};

typedef struct {
    int64_t nnz_off;        /* sum_k d_k  (off-diagonal entries of L)            */
    int64_t fills;          /* sum_k max(d_k - 1, 0)                            */
    int64_t gathered;       /* sum_k (entries gathered before merging)         */
    int64_t max_degree;     /* max_k d_k                                       */
    int32_t nlevels;        /* depth of the forward-solve DAG (levels 0..nl-1) */
    int32_t zero_pivots;    /* columns with D_k == 0                           */
} orc_factor_stats;

/* Randomized elimination (driver_local.cpp:302-707, gpu auxilliary.hpp:543-592).
 * Input: full symmetric Laplacian in CSC (both triangles; only the strict
 * lower part is read, as |value|).  Output: the off-diagonal part of L^T in
 * CSR form -- row k lists (r, -w_r/S_k) for r > k, sorted by r -- plus the
 * pivots diag[k] = S_k and the forward-solve level of every row.
 * lt_ind / lt_val need capacity `cap` (returns -1 if exceeded).
 * Returns 0 on success. */
int orc_factor(int n, const int64_t *colptr, const int *rowind, const double *val,
               uint64_t seed, int mutation, int64_t cap,
               int64_t *lt_ptr, int *lt_ind, double *lt_val, double *diag, int *level,
               orc_factor_stats *st);

/* Drop the ground vertex (last index) from the factor (driver_physics.cu:
 * 1945-2010): removes row n-1 of L^T and every entry with column n-1.
 * Returns the nnz written (n-1 rows remain). */
int64_t orc_factor_trim_last(int n, const int64_t *lt_ptr, const int *lt_ind, const double *lt_val,
                             int64_t *lt_ptr2, int *lt_ind2, double *lt_val2);

/* y = A x, CSR (== CSC for symmetric A), sequential fma per row. */
void orc_spmv(int n, const int64_t *ptr, const int *ind, const double *val, const double *x, double *y);

/* z = L^-T D+ L^-1 r with the reference's D+ rule (solver.hpp:63-84: divide
 * only where D != 0).  precision 64: all double; precision 32: factor values,
 * pivots and vectors in float, r rounded to float on entry.
 * Forward solve in push form over rows of L^T, backward in pull form, each
 * update acc = fma(-l, v, acc) in ascending index order. */
void orc_precond(int n, const int64_t *lt_ptr, const int *lt_ind, const double *lt_val,
                 const double *diag, int precision, int mutation, const double *r, double *z);

typedef struct {
    int32_t iterations;
    int32_t converged;
    double rz0;             /* r0' z0                                  */
    double rz;              /* final r' z                              */
    double rel_res_true;    /* ||A x - b|| / ||b||                     */
    double rel_res_rec;     /* ||r_k|| / ||b|| from the recurrence     */
} orc_pcg_stats;

enum { ORC_STOP_REFERENCE = 0, ORC_STOP_TRUE_RES = 1 };

/* PCG with the reference recurrence (solver.hpp:457-569), x0 = 0.
 * stop_rule REFERENCE: enter the loop only if r0'z0 > tol, continue while
 * sqrt(r'z)/sqrt(r0'z0) > tol and k < maxit.  TRUE_RES: continue while
 * ||r_k||/||b|| > tol.  Returns 0, or -3 on breakdown (p'Ap <= 0 or NaN). */
int orc_pcg(int n, const int64_t *a_ptr, const int *a_ind, const double *a_val,
            const int64_t *lt_ptr, const int *lt_ind, const double *lt_val, const double *diag,
            int precision, const double *b, double *x, double tol, int maxit, int stop_rule,
            orc_pcg_stats *st);

/* Incomplete Cholesky IC(0) L D L^T on the pattern of the lower triangle
 * (the context preconditioner; textbook recurrence, rows in index order):
 *   l_ij = (a_ij - sum_{m<j, (i,m),(j,m) in pattern} l_im l_jm d_m) / d_j,
 *   d_i  = a_ii - sum_j l_ij^2 d_j,
 * l_ij = 0 where d_j <= 0, d_i = 0 where it is not above 1e-14 a_ii.
 * Output in orc_factor's layout (off-diagonal L^T in CSR: row k lists the
 * rows r > k of column k of L with l_rk; pivots; forward levels).  Capacity
 * cap >= strict-lower nnz.  Returns 0, or -1 on capacity. */
int orc_ic0(int n, const int64_t *colptr, const int *rowind, const double *val, int mutation, int64_t cap,
            int64_t *lt_ptr, int *lt_ind, double *lt_val, double *diag, int *level);

/* ---- multi-threaded restatement (rchol_omp.c; CPU baseline) ---------------
 * orc_factor_omp: the same elimination as orc_factor, scheduled like the
 * reference's OpenMP factorization (driver_local.cpp:302-707, 802-829):
 * columns are taken from a shared ready queue as their lower neighbours
 * retire.  Bitwise equal to orc_factor for every thread count (threads <= 0:
 * all).  Returns 0, or -1 if the factor needs more than `cap` entries. */
int orc_factor_omp(int n, const int64_t *colptr, const int *rowind, const double *val, uint64_t seed,
                   int threads, int64_t cap, int64_t *lt_ptr, int *lt_ind, double *lt_val, double *diag,
                   int *level, orc_factor_stats *st);

/* orc_pcg_omp: the reference recurrence and stop rule (solver.hpp:457-569,
 * FP64) with OpenMP SpMV / dots / updates and level-scheduled triangular
 * solves over the factor's forward levels.  *setup_s = time to build the
 * level lists and rows of L. */
int orc_pcg_omp(int n, const int64_t *a_ptr, const int *a_ind, const double *a_val, const int64_t *lt_ptr,
                const int *lt_ind, const double *lt_val, const double *diag, const int *level, const double *b,
                double *x, double tol, int maxit, int threads, double *setup_s, orc_pcg_stats *st);

#ifdef __cplusplus
}
#endif
#endif
