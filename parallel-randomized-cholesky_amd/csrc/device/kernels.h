// kernels.h -- argument blocks and host-side launchers of the HIP kernels.
// Included by the host orchestration (capi.cpp) and by the .hip sources.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rc {

constexpr int kWaveRows256 = 256;   // threads of a sweep chunk group

// ------------------------------------------------------------ elimination
struct FillRec {            // one pending fill edge, 16 B
    int row;
    int next;               // next record of the same column, -1 = end
    double w;
};

// Device counter slots (unsigned long long each).
enum {
    kCtrFillTop = 2, kCtrOutTop = 3, kCtrStatus = 4,
    kCtrFills = 5, kCtrNnzOff = 8, kCtrMaxDeg = 7, kCtrZeroPivots = 10, kCtrScratchNeed = 11,
    kCtrCount = 16
};

// Device-side status word bits (ctr[kCtrStatus]).
enum : unsigned long long {
    kStatusOk = 0,
    kStatusFillPool = 1,
    kStatusOutPool = 2,
    kStatusScratch = 4,
    kStatusInconsistent = 16,
    kStatusTimeout = 32,
};

// Counters of the queue-driven tail (FactorArgs::aq_ctr), a cache line apart.
constexpr int kAqHead = 0, kAqTail = 32, kAqWords = 64;

// Pending fills of a column are kept in kFillLists singly linked lists
// (head[k * kFillLists + h]; a record goes to list (record index mod
// kFillLists)), so a column's gather walks them in parallel.
constexpr int kFillLists = 64;

// Columns with more than the round kernel's LDS capacity (128 entries) are
// deferred to a second kernel of the same round: kBigGrid one-wave workgroups
// with kBigCap entries of LDS each.  A round of at most kBigGrid columns is
// eliminated by that second kernel entirely.
constexpr int kBigGrid = 256;
constexpr int kBigCap = 2048;

// Ready lists are split into kListSegs segments of seg_cap entries (workgroup
// b appends to segment b % kListSegs) and an overflow region of n entries at
// the end: one shared append counter would serialize every append of a round
// (returning atomics on one address retire one at a time, ~11 ns apart on
// MI355X).  A round's counters are kListSlots words (the segments, then the
// overflow region); three sets rotate (round % 3).
constexpr int kListSegs = 63;
constexpr int kListSlots = kListSegs + 1;
inline long long list_seg_cap(int n) { return 2LL * ((n + kListSegs - 1) / kListSegs) + 1024; }
inline long long list_capacity(int n) { return (long long)kListSegs * list_seg_cap(n) + n; }

struct FactorArgs {
    int n;
    const int *lo_ptr;      // n+1, strict-lower part of the Laplacian (CSC)
    const int *lo_row;
    const double *lo_w;     // |A[r,k]|
    int *dep, *head, *level;   // head: n * kFillLists
    int *list0, *list1;     // ready lists of even / odd rounds (list_capacity(n) entries)
    unsigned int *seg_count;   // 3 x kListSlots append counters (set round % 3 = that round's list)
    int seg_cap;            // list_seg_cap(n)
    unsigned int *round_count;   // [round] = columns eliminated in that round (written by the round)
    long long *wave_state;  // per workgroup: fill cursor, fill end, out cursor, out end
    unsigned long long *ctr;
    FillRec *pool;
    long long pool_cap;
    int *out_row;
    double *out_val;
    long long out_cap;
    long long *out_start;
    int *out_cnt;
    double *diag;
    int *big_list;          // columns deferred in the current round
    unsigned int *big_count;   // [round] = deferred columns
    int big_ws;             // wave_state index of the first large-LDS workgroup
    char *scratch;          // overflow buffers of the kBigGrid deferred-column workgroups
    int scratch_cap;        // entries per workgroup (power of two)
    uint64_t seed;
    // queue-driven tail: column queue (n entries, -1 = not yet written) and
    // its head / tail counters; bounded waits
    int *aq;
    unsigned int *aq_ctr;
    unsigned int aq_total;  // slots of the queue in use (set by launch_factor_async)
    long long spin_ticks;
};

void launch_factor_setup(const FactorArgs &a, const int *dep_init, hipStream_t s);
// da: a device copy of a (the round and queue kernels read their arguments
// through it, so they are loaded when used instead of held in scalar
// registers for the whole kernel)
void launch_factor_rounds(const FactorArgs &a, const FactorArgs *da, int num_blocks, int first_round, int nrounds,
                          hipStream_t s);
void launch_factor_finish(const FactorArgs &a, hipStream_t s);
// The rest of the elimination from round `round` (whose list holds at most
// kBigGrid columns) in one persistent launch; `total` columns remain.
// da->aq_total must already hold `total` (the caller writes it into the device copy)
void launch_factor_async(const FactorArgs &a, const FactorArgs *da, int round, unsigned int total, hipStream_t s);
int factor_lds_capacity();
int factor_scratch_bytes_per_entry();
long long factor_pool_slack(int num_blocks);
int factor_occupancy_blocks(int device);
void factor_kernel_setup();   // large dynamic LDS of the deferred-column kernel

// ------------------------------------------------------------ assembly
// Exclusive scan of n int32 counts into out[0..n] (out[n] = total, so `out`
// needs n + 1 entries).  tmp needs scan_tmp_elems(n) ints and must not
// overlap `out`.
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

// Stable LSD radix sort of (uint32 key, int value) pairs; result in (k0, v0).
size_t radix_tmp_ints(int n);
void radix_sort_pairs(uint32_t *k0, int *v0, uint32_t *k1, int *v1, int n, int key_bits, int *tmp, hipStream_t s);

// Lt (row k = column k of L, off-diagonal, sorted) in precision T.
void launch_lt_counts(const AssembleArgs &a, int *cnt, hipStream_t s);
template <typename T>
void launch_lt_fill(const AssembleArgs &a, const int *lt_ptr, int *lt_ind, T *lt_val, T *diag_out, hipStream_t s);

// small helpers
void launch_iota(int *v, int n, hipStream_t s);
void launch_copy_keys(const int *src, uint32_t *dst, int n, hipStream_t s);
void launch_row_expand(int ns, const int *ptr, const int *map, int *rowk, hipStream_t s);
// keys[t] = idx_user[64 t] (user index of tile t's first position), vals[t] = t
void launch_tile_keys(int ntiles, const int *idx_user, uint32_t *keys, int *vals, hipStream_t s);
// ptr[r] = first index i with (keys[i] >> shift) >= r, r in [0, nseg]; keys sorted.
void launch_segment_bounds(int n, const uint32_t *keys, int shift, int nseg, int *ptr, hipStream_t s);
void launch_level_keys(int ns, const int *level, const int *lptr, int len_bits, uint32_t *keys, hipStream_t s);
void launch_inverse_perm(int n, const int *order, int *pos_of, hipStream_t s);
// keys of rows order[q]: level << 3 | width class (see assemble.hip)
void launch_level_bucket_keys(int ns, const int *order, const int *level, const int *l_ptr, const int *lt_ptr,
                              uint32_t *keys, hipStream_t s);
// keys of rows order[q]: level << 2 | sweep-group class (see assemble.hip)
void launch_level_group_keys(int ns, const int *order, const int *level, const int *l_ptr, const int *lt_ptr,
                             uint32_t *keys, hipStream_t s);
void launch_level_max(int ns, const int *level, unsigned long long *out, hipStream_t s);
// Rows per sweep chunk of a level whose widest row (of L and of L^T) has
// `width` entries: G = next power of two >= width / 8 lanes per row (gmin..64),
// so a lane holds at most eight entries of a row, and 256 / G rows (up to 256
// for gmin = 1: a level of short rows then keeps four times as many rows in
// flight per workgroup as with G = 4).
constexpr int kSweepMinGroupDefault = 1;
__host__ __device__ inline int chunk_rows(int width, int gmin)
{
    int g = gmin;
    while (g < 64 && g * 8 < width) g <<= 1;
    return kWaveRows256 / g;
}
// lw[L] = max over the rows of segment L of max(len L row, len L^T row)
// (lw zeroed by the caller); rows at positions p, segment of p =
// seg_keys[p] >> seg_shift (sorted keys) or, without keys, level[order[p]].
void launch_level_width(int ns, const int *order, const int *level, const uint32_t *seg_keys, int seg_shift,
                        const int *l_ptr, const int *lt_ptr, int *lw, hipStream_t s);
void launch_chunk_count(int nlevels, const int *lvl_ptr, const int *lw, int gmin, int *cnt, hipStream_t s);
void launch_chunk_fill(int nchunks, int nlevels, const int *lvl_ptr, const int *chunk_ptr, const int *lw, int gmin,
                       int *chunk_start, int *chunk_cnt, hipStream_t s);
// meta[c] = {start[c], cnt[c], off[c], slices[c]}
void launch_pack_meta(int nchunks, const int *start, const int *cnt, const int *off, const int *slices, int4 *meta,
                      hipStream_t s);
template <typename T> void launch_gather_t(const T *src, const int *idx, T *dst, int n, hipStream_t s);
void launch_gather_i(const int *src, const int *idx, int *dst, int n, hipStream_t s);
void launch_locality_keys(int ns, const int *perm, int nuser, const int *rank, uint32_t *keys, hipStream_t s);

// Sliced ELL (64 rows per slice, entry t of slice row l at off[c] + 64 t + l).
// Slice c covers rows order[start[c] + l], l < cnt[c] (chunk mode) or
// positions 64 c + l (tile mode, start == nullptr).  Entries come from CSR
// (ptr, ind, val) of the row (optionally through eidx: entry q of the row is
// src entry eidx[ptr[row] + q]) with the column mapped by colmap; padding is
// (pad_ind, pad_val).
struct EllSrc {
    const int *ptr;
    const int *ind;
    const void *val;      // T_src
    const int *eidx;      // optional indirection (nullptr = identity)
    const int *rowk;      // optional: column = rowk[e] instead of ind[e]
    const int *colmap;    // column -> position
};
void launch_ell_width(int nslices, const int *start, const int *cnt, const int *order, int nrows, const int *ptr,
                      int *width, hipStream_t s);
template <typename TS, typename TD>
void launch_ell_fill(int nslices, const int *start, const int *cnt, const int *order, int nrows, EllSrc src,
                     const int *off, const int *width, int pad_ind, TD pad_val, int *ell_ind, TD *ell_val,
                     hipStream_t s);
// off[c] = 64 * exclusive scan of width (int32); returns via scan kernels.
void launch_mul64(const int *in, int *out, int n, hipStream_t s);

// Grouped sliced ELL for the triangular sweeps: a chunk of cnt <= 256 rows is
// processed by one 256-thread workgroup with G = min(256 / P, 64) lanes per
// row (P = next power of two >= cnt).  Entry t of chunk row i lives in slice
// s = t / G at off[c] + s * P * G + i * G + (t % G).  slices[c] =
// ceil(width[c] / G); size[c] = slices[c] * P * G.
constexpr int kSweepThreads = 256;
constexpr int kTailThreads = 512;           // 8 waves: fewer idle waves per narrow level than 16
__host__ __device__ inline int sweep_rows_pow2(int cnt)
{
    return cnt <= 1 ? 1 : 1 << (32 - __builtin_clz((unsigned)(cnt - 1)));
}
__host__ __device__ inline int sweep_group(int cnt)
{
    int g = kSweepThreads / sweep_rows_pow2(cnt);
    return g > 64 ? 64 : g;
}
void launch_sweep_layout(int nchunks, const int *cnt, const int *width, int *slices, int *size, hipStream_t s);

// Tail of the level order: the rows at positions [tp0, ns) (the narrow
// trailing levels) are solved by one kTailThreads workgroup whose LDS holds
// their values.  The entries of each tail row's source row (a row of L for the
// forward sweep, of L^T for the backward sweep) are split by column position:
// below tp0 ("bulk part": final before the tail starts, folded in by a wide
// kernel) or at/after tp0 ("internal": read from LDS, column stored as
// position - tp0).  A tail level of cnt <= kTailMaxRows rows is one group
// with G = tail_group(cnt) lanes per row (P = next power of two >= cnt,
// lane slot i * G + g of the level's P * G) and slices = ceil(width / G)
// entries per lane slot, stored lane-major in blocks of kTailBlock: entry t
// of level row i (slice s = t / G, g = t % G) lives at
//   off[L] + (s / kTailBlock) * P * G * kTailBlock + (i * G + g) * kTailBlock + s % kTailBlock,
// so a lane loads a block's entries with a few 16-byte loads.  Unused entries
// are padding (index ntail, value 0); one more block of padding at off[ntl] is
// what lanes without a slot load.  The workgroup's LDS holds 0 at value slot
// ntail, so a padded contribution is -0 * 0 and needs no arithmetic mask.
constexpr int kTailMaxRows = kTailThreads;
constexpr int kTailBlock = 8;
__host__ __device__ inline int tail_group(int cnt)
{
    int g = kTailThreads / sweep_rows_pow2(cnt);
    return g > 64 ? 64 : (g < 1 ? 1 : g);
}
// LDS of the tail workgroup: ntail + 1 values (16-byte padded), then one
// int4 descriptor per tail level and one empty descriptor after them.
template <typename T>
__host__ __device__ inline int tail_value_bytes(int ntail)
{
    return (int)((((long long)ntail + 1) * (long long)sizeof(T) + 15) / 16 * 16);
}
void launch_tail_count(int ntail, int tp0, const int *order, EllSrc src, int *cnt_b, int *cnt_i, hipStream_t s);
template <typename TD>
void launch_tail_split(int ntail, int tp0, const int *order, EllSrc src, const int *ptr_b, const int *ptr_i,
                       int *ind_b, TD *val_b, int *ind_i, TD *val_i, hipStream_t s);
// slices[L] = ceil(max internal row length of level L / G),
// size[L] = ceil(slices / kTailBlock) * kTailBlock * P * G
void launch_tail_layout(int ntl, const int *lstart, const int *lcnt, int tp0, const int *ptr_i, int *slices,
                        int *size, hipStream_t s);
template <typename TD>
void launch_tail_fill(int ntl, const int *lstart, const int *lcnt, int tp0, const int *ptr_i, const int *ind_i,
                      const TD *val_i, const int *off, const int *slices, int pad_ind, int *ell_ind, TD *ell_val,
                      hipStream_t s);
template <typename TS, typename TD>
void launch_sweep_fill(int nchunks, const int *start, const int *cnt, const int *order, EllSrc src, const int *off,
                       const int *slices, int pad_ind, TD pad_val, int *ell_ind, TD *ell_val, hipStream_t s);

// ------------------------------------------------------------ PCG
struct PcgScalars {
    double rz, rz0, alpha, beta, pap, rr, bb;
    double tol;
    int iter, done, status, maxit, stop_rule, pad;
    unsigned int arrive_spmv, arrive_dot, arrive_res, pad2;
    double res2, bnorm2;
};
enum { kPcgOk = 0, kPcgBreakdown = 1, kPcgTimeout = 2 };

// Chunk claims of the persistent sweeps go to kClaimQueues counters
// kClaimStride words apart: chunk c belongs to queue c % K (K = min(queues,
// grid)) and workgroup b claims from queue b % K, so claims do not serialize
// on one address (returning atomics on one address retire ~11 ns apart).
// Forward counters first, then the backward ones.
constexpr int kClaimQueues = 16;
constexpr int kClaimStride = 64;
constexpr int kClaimWords = 2 * kClaimQueues * kClaimStride;
__device__ __forceinline__ void reset_claims(unsigned int *fwd, unsigned int *bwd, int t)
{
    if (t < kClaimQueues) {
        fwd[t * kClaimStride] = 0u;
        bwd[t * kClaimStride] = 0u;
    }
}

struct SolveArgs {
    int n;                        // rows; index n is a dummy slot holding 0
    // A (PCG system) in sliced ELL over position tiles
    int ntiles;
    const int *a_tile_ord;        // order in which the SpMVs visit the tiles (a permutation of 0..ntiles)
    const int *a_off, *a_width, *a_ind;
    const double *a_val;
    // level schedule (positions)
    int nchunks;
    const int *chunk_start, *chunk_cnt;
    const int4 *cm_f, *cm_b;      // per chunk {start, cnt, ELL offset, slices} of the two sweeps
    // L rows (forward) and L^T rows (backward) in grouped sliced ELL per chunk
    // (lf_width / lb_width hold the slice counts)
    const int *lf_off, *lf_width, *lf_ind;
    const int *lb_off, *lb_width, *lb_ind;
    const void *lf_val, *lb_val, *diag;
    // entry after the last chunk of each sweep's ELL: index n, value 0; what
    // lanes without a slot load when a chunk's slices are prefetched
    int lf_zero, lb_zero;
    // vectors in position numbering, n + 1 entries
    double *x, *r, *ap, *p0, *p1;
    void *y, *z;                  // precision of the factor, sentinel protocol
    PcgScalars *sc;
    double *partials;             // >= 2 * reduce blocks
    unsigned int *ctr_fwd, *ctr_bwd;
    int reduce_blocks;
    int sweep_blocks;             // persistent sweep workgroups (occupancy-sized)
    // level split: chunks [0, bulk_chunks) belong to the wide levels and run
    // on the persistent grid; levels [tail_level, nlevels) form the tail
    int bulk_chunks, tail_level, nlevels;
    // tail (see launch_tail_*): rows [tail_pos0, n), ntail_levels levels with
    // absolute start positions tl_start and row counts tl_cnt; internal ELL of
    // the forward (tf_*) and backward (tb_*) sweeps; CSR (tp_*) of the
    // forward bulk part (absolute column positions); tacc = bulk-folded
    // right-hand side of the tail rows (precision of the factor)
    int tail_pos0, ntail, ntail_levels, tail_lds_bytes;
    const int *tl_start, *tl_cnt;
    const int *tf_off, *tf_slices, *tf_ind;
    const int *tb_off, *tb_slices, *tb_ind;
    const void *tf_val, *tb_val;
    const int *tp_ptr, *tp_ind;
    const void *tp_val;
    void *tacc;
    long long spin_ticks;
    const double *b;
    int zero;                     // always 0 (keeps some loads per-lane, see k_sweep_bulk)
    int nap_max;                  // longest back-off between polls of a pending value (x s_sleep 2)
    // batched systems: system of every position and the per-system alpha
    // used by the forward sweep's r/x update (both null for one system)
    const int *sys_of;
    const double *alpha_sys;
};
int sweep_occupancy_blocks(int device);
// Largest dynamic LDS (bytes) the tail workgroup may use on this device.
int tail_lds_limit(int device);

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
// dst[0..n) = src[0..n), n even
void launch_copy(const double *src, double *dst, size_t n, hipStream_t s);

// ------------------------------------------------------------ ordering network
// Weight vector layout (float32): per layer Ws (64 x in), Wn (64 x in), b (64)
// for in = 2, 64, 64, then wo (64) and bo (1).
constexpr int kGnnHidden = 64;
constexpr int kGnnWs1 = 0, kGnnWn1 = 128, kGnnB1 = 256;
constexpr int kGnnWs2 = 320, kGnnWn2 = 320 + 4096, kGnnB2 = 320 + 8192;
constexpr int kGnnWs3 = kGnnB2 + 64, kGnnWn3 = kGnnWs3 + 4096, kGnnB3 = kGnnWs3 + 8192;
constexpr int kGnnWo = kGnnB3 + 64, kGnnBo = kGnnWo + 64;
constexpr int kGnnParams = kGnnBo + 1;   // 16897

struct GnnArgs {
    int n;
    const int *ptr, *ind;         // graph CSR (off-diagonal entries)
    const float *a;               // w / wdeg(row)        (forward aggregation)
    const float *at;              // w / wdeg(neighbour)  (transposed aggregation)
    float *H0, *M1;               // n x 2
    float *H[3], *M[3], *Z[3];    // n x 64 (index 0..2 = layers 1..3; M[0] unused)
    float *dA, *dB, *dC;          // n x 64 backward scratch
    float *part;                  // nblocks x (2*64*64 + 64)
    int nblocks;
};
// keep = true stores every layer (for the backward pass); otherwise layers
// ping-pong through H[0..1], M[0], Z[0].
void gnn_forward(const GnnArgs &g, const float *w, bool keep, float *scores, hipStream_t s);
// dw = d(sum_v ds[v] s[v]) / dw, using the stored forward pass.
void gnn_backward(const GnnArgs &g, const float *w, const float *ds, float *dw, hipStream_t s);

// ------------------------------------------------------------ multi-RHS PCG
constexpr int kMaxRhs = 64;
struct MultiScalars {
    double rz[kMaxRhs], rz0[kMaxRhs], alpha[kMaxRhs], beta[kMaxRhs], bb[kMaxRhs], res2[kMaxRhs];
    double rr[kMaxRhs];           // r'r of the last dot pass (true-residual stop rule)
    int iter[kMaxRhs], done[kMaxRhs], breakdown[kMaxRhs];
    double tol;
    int maxit, stop_rule, status, all_done;
    unsigned int arrive_spmv, arrive_dot, arrive_res, pad;
};

struct MultiArgs {
    int n, K, k;                  // rows, lane stride (pow2 >= k), right-hand sides
    const int *a_ptr, *a_ind;     // CSR in position numbering
    const double *a_val;
    const int *row_ord;           // position of every input-numbered row (the SpMV's visiting order)
    const int *l_ptr, *l_ind, *lt_ptr, *lt_ind;
    const void *l_val, *lt_val, *diag;
    int nchunks;
    const int *chunk_start, *chunk_cnt;
    double *x, *r, *ap, *p0, *p1;
    void *y, *z;
    const double *b;
    MultiScalars *sc;
    double *partials;             // 2 * reduce_blocks * K
    unsigned int *ctr_fwd, *ctr_bwd;
    int reduce_blocks, sweep_blocks;
    long long spin_ticks;
};

template <typename T> void launch_multi_init(const MultiArgs &a, hipStream_t s);
template <typename T> void launch_multi_iteration(const MultiArgs &a, int parity, hipStream_t s);
void launch_multi_residual(const MultiArgs &a, hipStream_t s);
// Single kernels of a multi-RHS iteration, for timing: the SpMM Ap = A p0
// (with its p'Ap partials), and the two sweeps (y, z sentinels first).
template <typename T> void launch_multi_spmm_only(const MultiArgs &a, hipStream_t s);
template <typename T> void launch_multi_fwd_only(const MultiArgs &a, hipStream_t s);
template <typename T> void launch_multi_bwd_only(const MultiArgs &a, hipStream_t s);
template <typename T> void launch_multi_sentinels(const MultiArgs &a, bool y, bool z, hipStream_t s);
void launch_multi_gather(const double *B, const int *idx_user, int n, int k, int K, double *b, hipStream_t s);
void launch_multi_scatter(const double *x, const int *idx_user, int n, int k, int K, double *X, hipStream_t s);

// ------------------------------------------------------------ batched systems
// nsys independent Laplacians stored as one block-diagonal system; every
// system runs its own PCG (own alpha, beta, r'z, stop test, iteration count).
// Reductions are segmented: positions of system s are listed (ascending) in
// spos[q], q in [qoff[s], qoff[s+1]); reduction block b covers
// q in [blk_lo[b], blk_hi[b]) of system blk_sys[b]; the blocks of system s are
// sys_blk0[s] .. sys_blk0[s+1]-1, and the last of them to finish reduces the
// block partials in block order (deterministic) and updates the scalars.
constexpr int kBatchRows = 2048;          // positions per reduction block
struct BatchArgs {
    int nsys, nblk;
    const int *sys_of;                    // position -> system
    const int *spos;                      // system-major positions
    const int *blk_lo, *blk_hi, *blk_sys, *sys_blk0;
    double *rz, *rz0, *alpha, *beta, *bb, *rr, *res2, *bn2;
    int *iter, *done, *brk;
    unsigned int *arrive;                 // per system
    unsigned int *arrive_all;             // systems finalised this pass
    double *part;                         // 2 * nblk
};
// Device bytes of the per-system state of nsys systems and nblk blocks.
size_t batch_state_bytes(int nsys, int nblk);
// Carve BatchArgs pointers out of one buffer of batch_state_bytes bytes.
void batch_carve(BatchArgs &b, void *buf);
template <typename T> void launch_batch_init(const SolveArgs &a, const BatchArgs &b, hipStream_t s);
template <typename T> void launch_batch_iteration(const SolveArgs &a, const BatchArgs &b, int parity, hipStream_t s);
void launch_batch_residual(const SolveArgs &a, const BatchArgs &b, hipStream_t s);
void launch_level_keys_sys(int ns, const int *level, const int *lptr, int len_bits, const int *sys, int sys_bits,
                           uint32_t *keys, hipStream_t s);

// ------------------------------------------------------------ input preparation (prep.hip)
void launch_map_rows(int nnz, const int *ind, const int *inv, uint32_t *rowk, hipStream_t s);
void launch_gather_key(const int *src, const int *idx, uint32_t *dst, int n, hipStream_t s);
void launch_perm_finish(int nnz, const int *idx, const int *ind, const int *inv, const double *val, int *out_ind,
                        double *out_val, hipStream_t s);
void launch_colsum(int n, const int *ptr, const double *val, double *cs, double *app, int *flag, hipStream_t s);
void launch_ground_count(int n, const int *ptr, const int *flag, int extra, int *cnt, hipStream_t s);
void launch_ground_fill(int n, const int *ptr, const int *ind, const double *val, const double *app, const int *flag,
                        const int *fpos, double corner, const int *optr, int *oind, double *oval, hipStream_t s);
void launch_trim_count(int n, const int *ptr, const int *ind, int *cnt, hipStream_t s);
void launch_trim_fill(int n, const int *ptr, const int *ind, const double *val, const int *optr, int *oind,
                      double *oval, hipStream_t s);
void launch_lower_count(int n, const int *ptr, const int *ind, int *cnt, int *dep, hipStream_t s);
void launch_lower_fill(int n, const int *ptr, const int *ind, const double *val, const int *lo_ptr, int *lo_row,
                       double *lo_w, hipStream_t s);
void launch_gumbel_keys(int n, const float *score, uint64_t seed, uint32_t *hi, uint32_t *lo, hipStream_t s);

// ------------------------------------------------------------ IC(0) (ic0.hip)
struct Ic0Args {
    int n;
    int n_compute;                // rows computed: n, or n - 1 when the last (ground) row is trimmed
    const int *up_ptr, *up_ind;   // row i of L: the rows j < i of column i of the prepared matrix
    const double *up_val, *diag_a;
    double *l;                    // values of L in the up_* layout
    double *d;                    // pivots, sentinel until published
    int *level;                   // forward-solve level of every row
    unsigned int *claim, *status; // chunk claim counter (zeroed), bit 0 = a wait expired
    long long spin_ticks;
};
void launch_ic0_upper(int n, const int *ptr, const int *ind, const double *val, const int *up_ptr, int *up_ind,
                      double *up_val, double *diag_a, hipStream_t s);
void launch_ic0(const Ic0Args &a, int waves, hipStream_t s);
void launch_ic0_scatter(int n, const int *up_ptr, const int *up_ind, const double *l, const int *lo_ptr,
                        const int *lo_row, double *out_val, hipStream_t s);
// out_start[k] = lo_ptr[k], out_cnt[k] = lo_ptr[k+1] - lo_ptr[k]
void launch_ptr_to_start_cnt(int n, const int *lo_ptr, long long *out_start, int *out_cnt, hipStream_t s);

// ------------------------------------------------------------ fixed workspace
// Device workspace of rc_solve_noalloc: a file-scope __device__ array of
// kNoAllocScratchBytes per GPU (noalloc.hip).  noalloc_workspace returns its
// device address on `device`, which must be the current device; the address
// is resolved once per device and cached.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
hipError_t noalloc_workspace(int device, void **addr);
// v[n] = 0 (one thread)
void launch_zero_slot(double *v, int n, hipStream_t s);

// CSR in position numbering: row p = source row order[p], columns mapped by
// src.colmap, values converted to TD.
void launch_pos_csr_count(int n, const int *order, const int *src_ptr, int *cnt, hipStream_t s);
template <typename TS, typename TD>
void launch_pos_csr_fill(int n, const int *order, EllSrc src, const int *ptr, int *ind, TD *val, hipStream_t s);

}  // namespace rc
