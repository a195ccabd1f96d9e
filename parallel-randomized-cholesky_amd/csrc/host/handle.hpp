// handle.hpp -- the state behind an rc_handle_t (one HIP stream and the
// device memory of one system) and the helpers the host sources of the C ABI
// share.  Internal: not part of the installed interface (include/rchol.h).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/rchol.h"
#include "../device/kernels.h"
#include "sparse.hpp"

using namespace rc;


struct DevVec {
    void *p = nullptr;
    size_t bytes = 0;
    ~DevVec() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    // grow-only; returns hipSuccess or the allocation error
    hipError_t ensure(size_t b)
    {
        if (b <= bytes && p) return hipSuccess;
        release();
        if (b == 0) b = 8;
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) { p = nullptr; return e; }
        bytes = b;
        return hipSuccess;
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

// Largest fill pool (records) and factor (off-diagonal entries) the int32
// indices of the fill lists and of the CSR of L^T can address.
constexpr long long kMaxPoolRecords = 2147483647LL;
constexpr long long kMaxFactorEntries = 2147483646LL;

inline int64_t pow2_at_least(int64_t v)
{
    int64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}


struct rc_handle_s {
    rc_options opt{};
    std::string err;
    hipStream_t stream = nullptr;
    long long spin_ticks = 0;
    int num_blocks = 0;
    bool host_only = false;

    // loaded system
    int kind = -1;
    Csc base;
    int n_user = 0;
    std::vector<int> perm;          // perm[new] = old (user numbering)
    int locality = 0;               // RC_LOCALITY_*: within-level row order of the sweeps
    std::vector<int> loc_rank;      // rank[user] in the recursive bisection order (empty: input order)
    double locality_s = 0.0;        // host seconds of the last bisection
    std::vector<float> gnn_w;

    // prepared system
    bool host_ready = false;
    bool prepared = false;
    Csc fac;                        // permuted Laplacian that is factored
    Csc sys;                        // permuted PCG system
    bool trim = false;
    int nf = 0, ns = 0;
    int64_t nnz_low = 0;

    // the loaded matrix on the device (b_*) and the device preparation's
    // buffers: permutation and inverse, factored matrix when it differs from
    // the PCG system (f_*), ground-append scratch (p_*)
    DevVec b_ptr, b_ind, b_val, pfull, pinv, f_ptr, f_ind, f_val, p_cs, p_app, p_flag, p_pos;
    bool base_dev = false;
    bool fac_in_f = false;          // the factored matrix is f_* (ground appended or kept), else a_*
    DevVec ic_up_ptr, ic_up_ind, ic_up_val, ic_l, ic_diag;   // IC(0) working set
    int64_t nnz_sys = 0;            // nnz of the PCG system
    double prepare_ms = 0.0;        // last preparation (ordering applied, lower slice, system)

    // factor inputs / working set
    DevVec lo_ptr, lo_row, lo_w, dep_init;
    DevVec big_list, big_count;
    DevVec dep, head, level, list0, list1, seg_count, round_count, wave_state, ctr, pool, out_row, out_val, out_start,
        out_cnt, diag64, scratch;
    DevVec aq, aq_ctr;              // queue of the elimination's queue-driven tail
    DevVec fargs;                   // device copy of the elimination's kernel arguments
    bool async_tail = true;         // narrow rounds run as one queue-driven launch
    int async_width = kBigGrid;     // ... once a round holds at most this many columns
    int sweep_gmin = kSweepMinGroupDefault;   // fewest lanes per row in a sweep chunk (chunks of <= 256 / gmin rows)
    long long pool_cap = 0, out_cap = 0;
    int scratch_cap = 0;
    int rounds = 0;

    // factor result: CSR of L^T in the factored numbering (export / dump)
    bool factored = false;
    int prec = 64;
    DevVec lt_ptr, lt_ind, lt_val, diagT;
    // solve layout: level-order positions, sliced ELL of L, L^T and A
    DevVec order, pos_of, idx_user, chunk_start, chunk_cnt, diagP, lvl_w, cm_f, cm_b;
    DevVec seg_work;                // chunk segment bounds, chunk counts and chunk offsets
    DevVec tile_ord, tile_k, tile_v, tile_tmp;   // SpMV tile order (by the user index of a tile's first row)
    DevVec pos_of_user;             // position of every user-numbered row
    DevVec lf_off, lf_width, lf_ind, lf_val, lb_off, lb_width, lb_ind, lb_val;
    DevVec a_off, a_width, a_ell_ind, a_ell_val;
    DevVec tmp_a, tmp_b, tmp_scan, keys0, keys1, vals0, vals1, rowk, l_ptr;
    int64_t nnz_lt = 0, ell_l = 0, ell_u = 0, ell_a = 0;
    int nlevels = 0, nchunks = 0, ntiles = 0;
    int tail_level = 0, bulk_chunks = 0;
    bool level_order_set = false;   // RCHOL_LEVEL_ORDER given
    int level_order = 1;            // 0 length, 1 spatial (default), 2 spatial within length buckets,
                                    // 3 spatial within sweep-group classes
    // tail of the level order (LDS-resident solve)
    int tail_max_rows = 256, tail_lds_cap = -1;
    int nap_max = 1;                // measured best: poll again after one s_sleep 2
    int seg_cap_limit = 0;          // > 0: ready-list segments hold at most this many entries (exercises the overflow region)
    int tail_pos0 = 0, ntail = 0, ntail_levels = 0, tail_lds_bytes = 0;
    DevVec tl_start, tl_cnt, tp_ptr, tp_ind, tp_val, tf_off, tf_slices, tf_ind, tf_val, tb_off, tb_slices, tb_ind,
        tb_val, ti_ptr, ti_ind, ti_val, tacc;
    rc_factor_stats fstats{};

    // PCG
    DevVec a_ptr, a_ind, a_val;
    DevVec x, r, ap, p0, p1, y, z, b, bu, scal, partials, ctrs, perm_d, loc_d, loc_pos;
    DevVec calib;                   // copy-kernel calibration buffer (rc_time_kernel)
    int reduce_blocks = 1024;
    int sweep_blocks = 1024;
    int sweep_blocks_max = 1024;    // occupancy limit of the persistent sweep grid
    bool rhs_set = false;
    std::vector<double> xref;
    int rhs_mode = 0;               // last resident right-hand side: 1 rc_set_rhs, 2 rc_set_rhs_multi
    // captured iteration pairs: graph (single / batch run), m_graph (multi-RHS
    // run); each is rebuilt when the buffers it was captured with change
    hipGraphExec_t graph = nullptr, m_graph = nullptr;
    const void *graph_key[3] = {nullptr, nullptr, nullptr};
    int graph_prec_of[3] = {0, 0, 0};
    int graph_K[3] = {0, 0, 0};

    // ordering network: graph + activations of the loaded system
    bool gnn_ready = false, gnn_keep = false;
    int gnn_n = 0;
    DevVec g_ptr, g_ind, g_a, g_at, g_H0, g_M1, g_H[3], g_M[3], g_Z[3], g_dA, g_dB, g_dC, g_part, g_w, g_dw, g_s, g_ds;

    // fixed workspace of rc_solve_noalloc on this handle's device (resolved at rc_create)
    void *na_ws = nullptr;
    hipError_t na_err = hipSuccess;

    // batched systems (nsys == 0: one system)
    int nsys = 0;
    std::vector<int64_t> sys_off;   // nsys + 1, user numbering
    std::vector<int> sys_user;      // user row -> system
    int nblk = 0;
    DevVec sys_user_d, sys_fac, sys_pos, spos, blk_tab, bt_state;
    std::vector<rc_solve_stats> bstats;

    // multi-RHS: CSRs in position numbering (built lazily per factor) + vectors
    bool multi_ready = false;
    DevVec m_aptr, m_aind, m_aval, m_lptr, m_lind, m_lval, m_ltptr, m_ltind, m_ltval;
    DevVec m_x, m_r, m_ap, m_p0, m_p1, m_y, m_z, m_b, m_B, m_X, m_sc, m_part;
    int m_k = 0;                    // right-hand sides held in m_B (user order, column-major)
    int m_n = 0;                    // rows of every right-hand side in m_B
    bool m_solved = false;          // m_x holds the solutions of the last multi-RHS run
    bool m_ran = false;             // a multi-RHS run sized the vectors for this factor (kernel timing)
    std::vector<rc_solve_stats> m_stats;   // per right-hand side of the last multi-RHS run

    ~rc_handle_s()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (m_graph) (void)hipGraphExecDestroy(m_graph);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

#define RC_HIP(expr)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            h->err = std::string(#expr) + ": " + hipGetErrorString(e_);                        \
            return e_ == hipErrorOutOfMemory ? RC_ERR_NOMEM : RC_ERR_HIP;                      \
        }                                                                                      \
    } while (0)

#define RC_HIP_H(hh, expr)                                                                     \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            (hh)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                     \
            return e_ == hipErrorOutOfMemory ? RC_ERR_NOMEM : RC_ERR_HIP;                      \
        }                                                                                      \
    } while (0)

#define RC_CHECK_HANDLE()                      \
    do {                                       \
        if (!h) return RC_ERR_ARG;             \
    } while (0)

inline int fail(rc_handle_t h, int code, const std::string &msg)
{
    h->err = msg;
    return code;
}

inline void invalidate_graph(rc_handle_t h)
{
    if (h->graph) (void)hipGraphExecDestroy(h->graph);
    if (h->m_graph) (void)hipGraphExecDestroy(h->m_graph);
    h->graph = nullptr;
    h->m_graph = nullptr;
}

// No C++ exception crosses the C ABI: host allocation failures become
// RC_ERR_NOMEM, any other exception RC_ERR_ARG, with the text in rc_last_error.
template <typename F>
int guarded(rc_handle_t h, F &&f) noexcept
{
    const char *what = "unknown exception";
    int code = RC_ERR_ARG;
    try {
        return f();
    } catch (const std::bad_alloc &) {
        what = "host memory allocation failed";
        code = RC_ERR_NOMEM;
    } catch (const std::length_error &) {
        what = "host container size limit exceeded";
        code = RC_ERR_NOMEM;
    } catch (const std::exception &e) {
        what = e.what();
    } catch (...) {
    }
    if (h) {
        try {
            h->err = what;
        } catch (...) {
        }
    }
    return code;
}

// Device preparation (prep_device.cpp).
int prepare_device(rc_handle_t h);
int learned_order_device(rc_handle_t h, const float *scores_dev, int n, uint64_t seed, std::vector<int> &perm);

// Shared between the host sources of the ABI (capi.cpp).
SolveArgs solve_args(rc_handle_t h);
