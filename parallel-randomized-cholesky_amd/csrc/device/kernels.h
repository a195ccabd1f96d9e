// kernels.h -- argument blocks and host-side launchers of the HIP kernels.
// Included by the host orchestration (capi.cpp) and by the .hip sources.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rc {

// ------------------------------------------------------------ elimination
struct FillRec {            // one pending fill edge, 16 B
    int row;
    int next;               // next record of the same column, -1 = end
    double w;
};

// Device counter slots (unsigned long long each).
enum {
    kCtrQHead = 0, kCtrQTail = 1, kCtrFillTop = 2, kCtrOutTop = 3, kCtrStatus = 4,
    kCtrFills = 5, kCtrGathered = 6, kCtrMaxDeg = 7, kCtrNnzOff = 8, kCtrMaxLevel = 9,
    kCtrZeroPivots = 10, kCtrScratchNeed = 11, kCtrCount = 16
};

// Device-side status word bits (ctr[kCtrStatus]).
enum : unsigned long long {
    kStatusOk = 0,
    kStatusFillPool = 1,
    kStatusOutPool = 2,
    kStatusScratch = 4,
    kStatusTimeout = 8,
    kStatusInconsistent = 16,
};

struct FactorArgs {
    int n;
    const int *lo_ptr;      // n+1, strict-lower part of the Laplacian (CSC)
    const int *lo_row;
    const double *lo_w;     // |A[r,k]|
    int *dep, *head, *nfill, *queue, *level;
    unsigned long long *ctr;
    FillRec *pool;
    long long pool_cap;
    int *out_row;
    double *out_val;
    long long out_cap;
    long long *out_start;
    int *out_cnt;
    double *diag;
    char *scratch;          // per-workgroup overflow buffers
    int scratch_cap;        // entries per workgroup (power of two)
    uint64_t seed;
    long long spin_ticks;
};

void launch_factor(const FactorArgs &a, int num_blocks, hipStream_t s);
int factor_lds_capacity();
int factor_scratch_bytes_per_entry();
int factor_occupancy_blocks(int device);

// ------------------------------------------------------------ assembly
// Exclusive scan of n int32 counts into out[0..n] (out[n] = total).
// tmp needs scan_tmp_elems(n) ints.
size_t scan_tmp_elems(int n);
void scan_exclusive(const int *in, int *out, int n, int *tmp, hipStream_t s);

struct AssembleArgs {
    int n;                  // factored columns
    int ns;                 // solve columns (n or n-1 when the ground is trimmed)
    int trim;               // drop vertex n-1
    const long long *out_start;
    const int *out_cnt;
    const int *out_row;
    const double *out_val;
    const double *diag_in;  // n
    const int *level;       // n
};

// Lt (row k = column k of L, off-diagonal, sorted) in precision T, then its
// transpose L (row r lists k < r ascending).  Pointers are int32.
void launch_lt_counts(const AssembleArgs &a, int *cnt, hipStream_t s);
template <typename T>
void launch_lt_fill(const AssembleArgs &a, const int *lt_ptr, int *lt_ind, T *lt_val, T *diag_out, hipStream_t s);
void launch_count_rows(int nnz, const int *ind, int *cnt, hipStream_t s);
template <typename T>
void launch_transpose_scatter(int ns, const int *lt_ptr, const int *lt_ind, const T *lt_val, int *pos,
                              int *l_ind, T *l_val, hipStream_t s);
template <typename T>
void launch_segmented_sort(int nseg, const int *ptr, int *ind, T *val, hipStream_t s);

// Level schedule: rows grouped by level, levels cut into chunks of <= 64 rows.
void launch_level_max(int ns, const int *level, unsigned long long *out, hipStream_t s);
void launch_level_count(int ns, const int *level, int *cnt, hipStream_t s);
void launch_level_scatter(int ns, const int *level, int *pos, int *order, hipStream_t s);
void launch_chunk_count(int nlevels, const int *lvl_ptr, int *cnt, hipStream_t s);
void launch_chunk_fill(int nchunks, int nlevels, const int *lvl_ptr, const int *chunk_ptr, int *chunk_start,
                       int *chunk_cnt, hipStream_t s);

// ------------------------------------------------------------ PCG
struct PcgScalars {
    double rz, rz0, alpha, beta, pap, rr, bb;
    double tol;
    int iter, done, status, maxit, stop_rule, pad;
    unsigned int arrive_spmv, arrive_dot, arrive_res, pad2;
    double res2, bnorm2;
};
enum { kPcgOk = 0, kPcgBreakdown = 1, kPcgTimeout = 2 };

struct SolveArgs {
    int n;
    // A (PCG system), CSR == CSC
    const int *a_ptr;
    const int *a_ind;
    const double *a_val;
    // factor
    int nchunks;
    const int *chunk_start, *chunk_cnt, *order;
    const int *l_ptr, *l_ind;     // rows of L (forward, pull)
    const int *lt_ptr, *lt_ind;   // rows of L^T (backward, pull)
    const void *l_val, *lt_val, *diag;
    // vectors
    double *x, *r, *ap, *p0, *p1;
    void *y, *z;                  // precision of the factor, sentinel protocol
    PcgScalars *sc;
    double *partials;             // >= reduce blocks
    unsigned int *ctr_fwd, *ctr_bwd;
    int reduce_blocks;
    long long spin_ticks;
    const double *b;
};

template <typename T> void launch_fill_sentinel(T *v, int n, hipStream_t s);
template <typename T> void launch_pcg_init(const SolveArgs &a, hipStream_t s);
template <typename T> void launch_pcg_iteration(const SolveArgs &a, int parity, hipStream_t s);
template <typename T> void launch_spmv_only(const SolveArgs &a, hipStream_t s);
template <typename T> void launch_trsv_fwd_only(const SolveArgs &a, hipStream_t s);
template <typename T> void launch_trsv_bwd_only(const SolveArgs &a, hipStream_t s);
void launch_residual(const SolveArgs &a, hipStream_t s);
void launch_gather(const double *src, const int *idx, double *dst, int n, hipStream_t s);
void launch_scatter(const double *src, const int *idx, double *dst, int n, hipStream_t s);
void launch_reset_counters(unsigned int *a, unsigned int *b, hipStream_t s);

}  // namespace rc
