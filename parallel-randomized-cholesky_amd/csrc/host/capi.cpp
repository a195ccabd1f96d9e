// capi.cpp -- implementation of the rchol.h C ABI: one handle = one HIP
// stream + the device memory of one system.  Host work is limited to the
// input boundary (Matrix Market, Laplacian build, ordering, ground append);
// the factorization and the PCG run in the HIP kernels of csrc/device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <stdexcept>
#include <random>
#include <string>
#include <vector>

#include "../../../include/rchol.h"
#include "../device/kernels.h"
#include "handle.hpp"
#include "mmio.hpp"
#include "philox.h"
#include "sparse.hpp"
#include "gnn_host.hpp"

using namespace rc;

extern "C" {

const char *rc_status_string(int s)
{
    switch (s) {
    case RC_OK: return "ok";
    case RC_ERR_ARG: return "bad argument";
    case RC_ERR_IO: return "i/o error";
    case RC_ERR_FORMAT: return "malformed input";
    case RC_ERR_NOT_DOMINANT: return "not diagonally dominant";
    case RC_ERR_STATE: return "call out of order";
    case RC_ERR_HIP: return "HIP runtime error";
    case RC_ERR_POOL: return "device pool exhausted";
    case RC_ERR_TIMEOUT: return "device wait timed out";
    case RC_ERR_BREAKDOWN: return "PCG breakdown";
    case RC_ERR_NO_DEVICE: return "no HIP device";
    case RC_ERR_NOMEM: return "out of memory";
    case RC_ERR_NOT_SYMMETRIC: return "matrix not symmetric";
    case RC_ERR_MAXIT: return "maximum iterations reached";
    default: return "unknown status";
    }
}

void rc_default_options(rc_options *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->device = 0;
    o->seed = 1;
    o->num_blocks = 0;
    o->factor_precision = 64;
    o->stop_rule = RC_STOP_REFERENCE;
    o->maxit = 300;
    o->pool_factor = 4.0;
    o->verbose = 0;
}

static int rc_create_impl(rc_handle_t *out, const rc_options *opt)
{
    if (!out) return RC_ERR_ARG;
    *out = nullptr;
    rc_options o;
    rc_default_options(&o);
    if (opt) o = *opt;
    if (o.factor_precision != 32 && o.factor_precision != 64) return RC_ERR_ARG;
    if (o.maxit < 0 || o.pool_factor <= 0.0 || o.host_threads < 0) return RC_ERR_ARG;
    if (o.device == -1) {                           // host-only handle: input, ordering, RHS
        rc_handle_t h = new (std::nothrow) rc_handle_s();
        if (!h) return RC_ERR_NOMEM;
        h->opt = o;
        h->host_only = true;
        *out = h;
        return RC_OK;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return RC_ERR_NO_DEVICE;
    if (o.device < 0 || o.device >= ndev) return RC_ERR_ARG;
    rc_handle_t h = new (std::nothrow) rc_handle_s();
    if (!h) return RC_ERR_NOMEM;
    h->opt = o;
    if (hipSetDevice(o.device) != hipSuccess || hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return RC_ERR_HIP;
    }
    int rate_khz = 0;
    if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, o.device) != hipSuccess || rate_khz <= 0)
        rate_khz = 100000;
    h->spin_ticks = (long long)rate_khz * 1000LL * 20LL;        // 20 s
    // device address of the fixed workspace of rc_solve_noalloc, resolved once
    // per device with the device current (an error is reported by that entry point)
    h->na_err = noalloc_workspace(o.device, &h->na_ws);
    h->num_blocks = o.num_blocks > 0 ? o.num_blocks : factor_occupancy_blocks(o.device);
    factor_kernel_setup();
    h->sweep_blocks = sweep_occupancy_blocks(o.device);
    h->sweep_blocks_max = h->sweep_blocks;
    {
        const char *lo = std::getenv("RCHOL_LEVEL_ORDER");      // "spatial" (default) | "length" | "bucket"
        if (lo && std::string(lo) == "length") h->level_order = 0;
        if (lo && std::string(lo) == "spatial") h->level_order = 1;
        if (lo && std::string(lo) == "bucket") h->level_order = 2;
        if (lo && std::string(lo) == "groups") h->level_order = 3;
        if (lo) h->level_order_set = true;
        const char *nm = std::getenv("RCHOL_NAP_MAX");          // polling back-off cap (1..64)
        if (nm) h->nap_max = std::max(1, std::min(64, std::atoi(nm)));
        const char *tr = std::getenv("RCHOL_TAIL_ROWS");        // tail level width limit (1..kTailMaxRows + 1)
        if (tr) h->tail_max_rows = std::max(1, std::min(kTailMaxRows + 1, std::atoi(tr)));
        const char *sc = std::getenv("RCHOL_LIST_SEG_CAP");     // ready-list segment capacity limit (>= 1)
        if (sc) h->seg_cap_limit = std::max(1, std::atoi(sc));
        const char *at = std::getenv("RCHOL_ASYNC_TAIL");       // 0: narrow rounds stay round kernels
        if (at) h->async_tail = std::atoi(at) != 0;
        const char *aw = std::getenv("RCHOL_ASYNC_WIDTH");      // round width at which the queue-driven tail starts
        if (aw) h->async_width = std::max(1, std::atoi(aw));
        const char *mg = std::getenv("RCHOL_SWEEP_MIN_GROUP");  // lanes per row of the widest sweep chunks (1, 2, 4)
        if (mg) {
            const int v = std::atoi(mg);
            h->sweep_gmin = v <= 1 ? 1 : (v <= 2 ? 2 : 4);
        }
    }
    *out = h;
    return RC_OK;
}

static int rc_destroy_impl(rc_handle_t h)
{
    if (!h) return RC_ERR_ARG;
    if (!h->host_only) (void)hipSetDevice(h->opt.device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    delete h;
    return RC_OK;
}

const char *rc_last_error(rc_handle_t h)
{
    if (!h) return "null handle";
    return h->err.c_str();
}

// ------------------------------------------------------------------ input
static void reset_system(rc_handle_t h)
{
    h->gnn_ready = false;
    h->host_ready = false;
    h->prepared = false;
    h->factored = false;
    h->rhs_set = false;
    h->perm.clear();
    h->loc_rank.clear();
    h->xref.clear();
    h->nsys = 0;
    h->sys_off.clear();
    h->sys_user.clear();
    h->bstats.clear();
    // multi-RHS state belongs to the previous system: its vectors were sized
    // for that system's n
    h->base_dev = false;
    h->multi_ready = false;
    h->m_k = 0;
    h->m_n = 0;
    h->m_solved = false;
    h->m_ran = false;
    // pools are sized from the new system's nnz at its first factor (the
    // device buffers themselves are grow-only and stay allocated)
    h->pool_cap = 0;
    h->out_cap = 0;
    h->scratch_cap = 0;
    invalidate_graph(h);
}

static int rc_set_triplets_impl(rc_handle_t h, int n, int64_t m, const int *rows, const int *cols, const double *vals, int kind)
{
    RC_CHECK_HANDLE();
    if (n <= 0 || m < 0 || (m > 0 && (!rows || !cols || !vals))) return fail(h, RC_ERR_ARG, "bad triplet arrays");
    if (kind != RC_INPUT_LAPLACIAN && kind != RC_INPUT_GROUNDED && kind != RC_INPUT_SDDM)
        return fail(h, RC_ERR_ARG, "unknown input kind");
    reset_system(h);
    Csc a;
    int rc = build_laplacian(n, m, rows, cols, vals, a);
    if (rc != RC_OK) return fail(h, rc, "triplet index out of range");
    if (!is_symmetric(a)) return fail(h, RC_ERR_NOT_SYMMETRIC, "input matrix is not symmetric");
    if (kind == RC_INPUT_GROUNDED && n < 2) return fail(h, RC_ERR_ARG, "grounded input needs n >= 2");
    h->base = std::move(a);
    h->kind = kind;
    h->n_user = kind == RC_INPUT_GROUNDED ? n - 1 : n;
    return RC_OK;
}

static int rc_set_batch_impl(rc_handle_t h, int nsys, const int *sys_n, const int64_t *sys_m, const int *rows, const int *cols,
                 const double *vals)
{
    RC_CHECK_HANDLE();
    if (nsys <= 0 || !sys_n || !sys_m) return fail(h, RC_ERR_ARG, "bad batch description");
    int64_t n = 0, m = 0;
    for (int s = 0; s < nsys; ++s) {
        if (sys_n[s] <= 0 || sys_m[s] < 0) return fail(h, RC_ERR_ARG, "every system needs n >= 1 and m >= 0");
        n += sys_n[s];
        m += sys_m[s];
    }
    if (n > 2147483646LL) return fail(h, RC_ERR_ARG, "batch exceeds int32 rows");
    if (m > 0 && (!rows || !cols || !vals)) return fail(h, RC_ERR_ARG, "bad triplet arrays");
    // system-local indices -> block-diagonal numbering
    std::vector<int> r((size_t)m), c((size_t)m);
    int64_t q = 0, off = 0;
    for (int s = 0; s < nsys; ++s) {
        for (int64_t e = 0; e < sys_m[s]; ++e, ++q) {
            if (rows[q] < 0 || rows[q] >= sys_n[s] || cols[q] < 0 || cols[q] >= sys_n[s])
                return fail(h, RC_ERR_ARG, "triplet index outside its system (system " + std::to_string(s) + ")");
            r[(size_t)q] = (int)(rows[q] + off);
            c[(size_t)q] = (int)(cols[q] + off);
        }
        off += sys_n[s];
    }
    int rc = rc_set_triplets(h, (int)n, m, r.data(), c.data(), vals, RC_INPUT_LAPLACIAN);
    if (rc != RC_OK) return rc;
    h->nsys = nsys;
    h->sys_off.assign((size_t)nsys + 1, 0);
    h->sys_user.resize((size_t)n);
    for (int s = 0; s < nsys; ++s) {
        h->sys_off[(size_t)s + 1] = h->sys_off[(size_t)s] + sys_n[s];
        for (int64_t i = h->sys_off[(size_t)s]; i < h->sys_off[(size_t)s + 1]; ++i) h->sys_user[(size_t)i] = s;
    }
    return RC_OK;
}

static int rc_get_batch_impl(rc_handle_t h, int *nsys, int64_t *offsets)
{
    RC_CHECK_HANDLE();
    if (h->kind < 0) return fail(h, RC_ERR_STATE, "no matrix loaded");
    if (nsys) *nsys = h->nsys > 0 ? h->nsys : 1;
    if (offsets) {
        if (h->nsys > 0) std::memcpy(offsets, h->sys_off.data(), h->sys_off.size() * sizeof(int64_t));
        else { offsets[0] = 0; offsets[1] = h->n_user; }
    }
    return RC_OK;
}

static int rc_load_mtx_impl(rc_handle_t h, const char *path, int kind)
{
    RC_CHECK_HANDLE();
    if (!path) return fail(h, RC_ERR_ARG, "null path");
    MtxTriplets t;
    int rc = read_mtx(path, t, h->err, h->opt.host_threads);
    if (rc != RC_OK) return rc;
    if (t.nrows != t.ncols) return fail(h, RC_ERR_FORMAT, "matrix is not square");   // pre_process.hpp:358
    return rc_set_triplets(h, t.nrows, (int64_t)t.rows.size(), t.rows.data(), t.cols.data(), t.vals.data(), kind);
}

static int rc_get_size_impl(rc_handle_t h, int64_t *n, int64_t *nnz)
{
    RC_CHECK_HANDLE();
    if (h->kind < 0) return fail(h, RC_ERR_STATE, "no matrix loaded");
    if (n) *n = h->n_user;
    if (nnz) {
        if (h->kind == RC_INPUT_GROUNDED) {
            Csc t;
            trim_last(h->base, t);
            *nnz = t.nnz();
        } else {
            *nnz = h->base.nnz();
        }
    }
    return RC_OK;
}

// ------------------------------------------------------------------ ordering
static int gnn_setup(rc_handle_t h, bool keep);
static GnnArgs gnn_args(rc_handle_t h);

// The recursive bisection order of the loaded graph as ranks per user vertex
// (once per matrix; the device copy is made at preparation).
static void compute_locality(rc_handle_t h)
{
    h->loc_rank.clear();
    if (h->locality != RC_LOCALITY_BISECTION || h->kind < 0 || h->n_user <= 0) return;
    const auto t0 = std::chrono::steady_clock::now();
    h->loc_rank.assign((size_t)h->n_user, 0);
    if (h->nsys == 0) {
        const std::vector<int> p = rb_perm(h->base, h->n_user, 1, h->opt.host_threads);
        for (int i = 0; i < h->n_user; ++i) h->loc_rank[(size_t)p[(size_t)i]] = i;
    } else {
        // a batch: each system bisected on its own, ranks kept inside the
        // system's row range (the systems stay contiguous and in order)
        for (int q = 0; q < h->nsys; ++q) {
            const int64_t o = h->sys_off[(size_t)q], e = h->sys_off[(size_t)q + 1];
            Csc sub;
            sub.n = (int)(e - o);
            sub.ptr.assign((size_t)sub.n + 1, 0);
            for (int64_t j = o; j < e; ++j) {
                for (int64_t k = h->base.ptr[(size_t)j]; k < h->base.ptr[(size_t)j + 1]; ++k) {
                    const int i = h->base.ind[(size_t)k];
                    if (i < o || i >= e) continue;
                    sub.ind.push_back((int)(i - o));
                    sub.val.push_back(h->base.val[(size_t)k]);
                }
                sub.ptr[(size_t)(j - o) + 1] = (int64_t)sub.ind.size();
            }
            const std::vector<int> p = rb_perm(sub, sub.n, 1, h->opt.host_threads);
            for (int i = 0; i < sub.n; ++i) h->loc_rank[(size_t)(o + p[(size_t)i])] = (int)(o + i);
        }
    }
    h->locality_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

static int rc_set_locality_impl(rc_handle_t h, int mode)
{
    RC_CHECK_HANDLE();
    if (mode != RC_LOCALITY_INPUT && mode != RC_LOCALITY_BISECTION) return fail(h, RC_ERR_ARG, "unknown locality mode");
    if (mode == h->locality && (mode == RC_LOCALITY_INPUT || !h->loc_rank.empty() || h->kind < 0)) return RC_OK;
    h->locality = mode;
    compute_locality(h);
    // the level order is built at preparation/assembly: prepare again
    h->prepared = false;
    h->factored = false;
    h->rhs_set = false;
    invalidate_graph(h);
    return RC_OK;
}

static int rc_get_locality_impl(rc_handle_t h, int *rank, double *seconds)
{
    RC_CHECK_HANDLE();
    if (h->kind < 0 || !rank) return fail(h, RC_ERR_ARG, "no matrix or null array");
    for (int i = 0; i < h->n_user; ++i) rank[i] = h->loc_rank.empty() ? i : h->loc_rank[(size_t)i];
    if (seconds) *seconds = h->locality_s;
    return RC_OK;
}

static int rc_order_impl(rc_handle_t h, int method, uint64_t seed)
{
    RC_CHECK_HANDLE();
    if (h->kind < 0) return fail(h, RC_ERR_STATE, "no matrix loaded");
    const int nv = h->n_user;
    const int skip = h->kind == RC_INPUT_GROUNDED ? h->base.n - 1 : -1;
    std::vector<int> p;
    switch (method) {
    case RC_ORDER_AS_IS:
        p.resize((size_t)nv);
        for (int i = 0; i < nv; ++i) p[(size_t)i] = i;
        break;
    case RC_ORDER_NNZ_SORT: p = nnz_sort_perm(h->base, nv, seed, skip); break;
    case RC_ORDER_RANDOM: p = random_perm(nv, seed); break;
    case RC_ORDER_AMD:
        p = amd_perm(h->base, nv);
        if ((int)p.size() != nv) return fail(h, RC_ERR_ARG, "AMD ordering failed");
        break;
    case RC_ORDER_ND: p = nd_perm(h->base, nv, seed, h->opt.host_threads); break;
    case RC_ORDER_LEARNED:
        if (h->gnn_w.empty() || h->host_only) p = nnz_sort_perm(h->base, nv, seed, skip);      // fallback
        else {
            // network scores and the Plackett-Luce draw stay on the device
            RC_HIP(hipSetDevice(h->opt.device));
            int rc = gnn_setup(h, false);
            if (rc != RC_OK) return rc;
            RC_HIP(hipMemcpyAsync(h->g_w.p, h->gnn_w.data(), (size_t)kGnnParams * sizeof(float), hipMemcpyHostToDevice,
                                  h->stream));
            GnnArgs g = gnn_args(h);
            gnn_forward(g, h->g_w.as<float>(), h->gnn_keep, h->g_s.as<float>(), h->stream);
            RC_HIP(hipGetLastError());
            rc = learned_order_device(h, h->g_s.as<float>(), nv, seed, p);
            if (rc != RC_OK) return rc;
        }
        break;
    default: return fail(h, RC_ERR_ARG, "unknown ordering");
    }
    h->perm = std::move(p);
    if (h->locality == RC_LOCALITY_BISECTION && h->loc_rank.empty()) compute_locality(h);
    h->host_ready = false;
    h->prepared = false;
    h->factored = false;
    h->rhs_set = false;
    invalidate_graph(h);
    return RC_OK;
}

static int rc_get_permutation_impl(rc_handle_t h, int *perm)
{
    RC_CHECK_HANDLE();
    if (h->kind < 0 || !perm) return fail(h, RC_ERR_ARG, "no matrix or null array");
    for (int i = 0; i < h->n_user; ++i) perm[i] = h->perm.empty() ? i : h->perm[(size_t)i];
    return RC_OK;
}

static int rc_set_gnn_weights_impl(rc_handle_t h, const float *w, int64_t count)
{
    RC_CHECK_HANDLE();
    if (count < 0 || (count > 0 && !w)) return fail(h, RC_ERR_ARG, "bad weights");
    if (count > 0 && count != gnn_weight_count()) return fail(h, RC_ERR_ARG, "weight count mismatch");
    h->gnn_w.assign(w, w + count);
    return RC_OK;
}

static int rc_set_permutation_impl(rc_handle_t h, const int *perm)
{
    RC_CHECK_HANDLE();
    if (h->kind < 0) return fail(h, RC_ERR_STATE, "no matrix loaded");
    if (!perm) return fail(h, RC_ERR_ARG, "null permutation");
    const int nv = h->n_user;
    std::vector<char> seen((size_t)nv, 0);
    for (int i = 0; i < nv; ++i) {
        if (perm[i] < 0 || perm[i] >= nv || seen[(size_t)perm[i]]) return fail(h, RC_ERR_ARG, "not a permutation");
        seen[(size_t)perm[i]] = 1;
    }
    h->perm.assign(perm, perm + nv);
    if (h->locality == RC_LOCALITY_BISECTION && h->loc_rank.empty()) compute_locality(h);
    h->host_ready = false;
    h->prepared = false;
    h->factored = false;
    h->rhs_set = false;
    invalidate_graph(h);
    return RC_OK;
}

int64_t rc_gnn_weight_count(void) { return gnn_weight_count(); }

static int gnn_setup(rc_handle_t h, bool keep)
{
    if (h->gnn_ready && (h->gnn_keep || !keep)) return RC_OK;
    GnnGraph g;
    gnn_graph(h->base, h->n_user, g);
    const size_t n = (size_t)g.n, m = std::max<size_t>(g.ind.size(), 1);
    hipStream_t s = h->stream;
    RC_HIP(h->g_ptr.ensure((n + 1) * sizeof(int)));
    RC_HIP(h->g_ind.ensure(m * sizeof(int)));
    RC_HIP(h->g_a.ensure(m * sizeof(float)));
    RC_HIP(h->g_at.ensure(m * sizeof(float)));
    RC_HIP(h->g_H0.ensure(n * 2 * sizeof(float) + 16));
    RC_HIP(h->g_M1.ensure(n * 2 * sizeof(float) + 16));
    const int nl = keep ? 3 : 2;
    for (int l = 0; l < nl; ++l) RC_HIP(h->g_H[l].ensure(n * kGnnHidden * sizeof(float) + 16));
    for (int l = 0; l < (keep ? 3 : 1); ++l) {
        RC_HIP(h->g_M[l].ensure(n * kGnnHidden * sizeof(float) + 16));
        RC_HIP(h->g_Z[l].ensure(n * kGnnHidden * sizeof(float) + 16));
    }
    if (keep) {
        RC_HIP(h->g_dA.ensure(n * kGnnHidden * sizeof(float) + 16));
        RC_HIP(h->g_dB.ensure(n * kGnnHidden * sizeof(float) + 16));
        RC_HIP(h->g_dC.ensure(n * kGnnHidden * sizeof(float) + 16));
        RC_HIP(h->g_part.ensure((size_t)256 * (2 * kGnnHidden * kGnnHidden + kGnnHidden) * sizeof(float)));
        RC_HIP(h->g_dw.ensure((size_t)kGnnParams * sizeof(float)));
        RC_HIP(h->g_ds.ensure(n * sizeof(float) + 16));
    }
    RC_HIP(h->g_w.ensure((size_t)kGnnParams * sizeof(float)));
    RC_HIP(h->g_s.ensure(n * sizeof(float) + 16));
    RC_HIP(hipMemcpyAsync(h->g_ptr.p, g.ptr.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice, s));
    if (!g.ind.empty()) {
        RC_HIP(hipMemcpyAsync(h->g_ind.p, g.ind.data(), g.ind.size() * sizeof(int), hipMemcpyHostToDevice, s));
        RC_HIP(hipMemcpyAsync(h->g_a.p, g.a.data(), g.a.size() * sizeof(float), hipMemcpyHostToDevice, s));
        RC_HIP(hipMemcpyAsync(h->g_at.p, g.at.data(), g.at.size() * sizeof(float), hipMemcpyHostToDevice, s));
    }
    RC_HIP(hipMemcpyAsync(h->g_H0.p, g.x0.data(), n * 2 * sizeof(float), hipMemcpyHostToDevice, s));
    RC_HIP(hipStreamSynchronize(s));
    h->gnn_n = g.n;
    h->gnn_ready = true;
    h->gnn_keep = keep;
    return RC_OK;
}

static GnnArgs gnn_args(rc_handle_t h)
{
    GnnArgs g;
    g.n = h->gnn_n;
    g.ptr = h->g_ptr.as<int>();
    g.ind = h->g_ind.as<int>();
    g.a = h->g_a.as<float>();
    g.at = h->g_at.as<float>();
    g.H0 = h->g_H0.as<float>();
    g.M1 = h->g_M1.as<float>();
    for (int l = 0; l < 3; ++l) {
        g.H[l] = h->g_H[l].as<float>();
        g.M[l] = h->g_M[l].as<float>();
        g.Z[l] = h->g_Z[l].as<float>();
    }
    g.dA = h->g_dA.as<float>();
    g.dB = h->g_dB.as<float>();
    g.dC = h->g_dC.as<float>();
    g.part = h->g_part.as<float>();
    g.nblocks = 256;
    return g;
}

static int rc_gnn_scores_impl(rc_handle_t h, const float *w, float *scores)
{
    RC_CHECK_HANDLE();
    if (h->host_only) return fail(h, RC_ERR_NO_DEVICE, "host-only handle");
    if (h->kind < 0) return fail(h, RC_ERR_STATE, "no matrix loaded");
    if (!w || !scores) return fail(h, RC_ERR_ARG, "null array");
    RC_HIP(hipSetDevice(h->opt.device));
    int rc = gnn_setup(h, false);
    if (rc != RC_OK) return rc;
    hipStream_t s = h->stream;
    RC_HIP(hipMemcpyAsync(h->g_w.p, w, (size_t)kGnnParams * sizeof(float), hipMemcpyHostToDevice, s));
    GnnArgs g = gnn_args(h);
    gnn_forward(g, h->g_w.as<float>(), h->gnn_keep, h->g_s.as<float>(), s);
    RC_HIP(hipGetLastError());
    RC_HIP(hipMemcpyAsync(scores, h->g_s.p, (size_t)h->gnn_n * sizeof(float), hipMemcpyDeviceToHost, s));
    RC_HIP(hipStreamSynchronize(s));
    return RC_OK;
}

static int rc_gnn_gradient_impl(rc_handle_t h, const float *w, const float *dscores, float *dw)
{
    RC_CHECK_HANDLE();
    if (h->host_only) return fail(h, RC_ERR_NO_DEVICE, "host-only handle");
    if (h->kind < 0) return fail(h, RC_ERR_STATE, "no matrix loaded");
    if (!w || !dscores || !dw) return fail(h, RC_ERR_ARG, "null array");
    RC_HIP(hipSetDevice(h->opt.device));
    int rc = gnn_setup(h, true);
    if (rc != RC_OK) return rc;
    hipStream_t s = h->stream;
    RC_HIP(hipMemcpyAsync(h->g_w.p, w, (size_t)kGnnParams * sizeof(float), hipMemcpyHostToDevice, s));
    RC_HIP(hipMemcpyAsync(h->g_ds.p, dscores, (size_t)h->gnn_n * sizeof(float), hipMemcpyHostToDevice, s));
    GnnArgs g = gnn_args(h);
    gnn_forward(g, h->g_w.as<float>(), true, h->g_s.as<float>(), s);
    RC_HIP(hipMemsetAsync(h->g_dw.p, 0, (size_t)kGnnParams * sizeof(float), s));
    gnn_backward(g, h->g_w.as<float>(), h->g_ds.as<float>(), h->g_dw.as<float>(), s);
    RC_HIP(hipGetLastError());
    RC_HIP(hipMemcpyAsync(dw, h->g_dw.p, (size_t)kGnnParams * sizeof(float), hipMemcpyDeviceToHost, s));
    RC_HIP(hipStreamSynchronize(s));
    return RC_OK;
}

// ------------------------------------------------------------------ prepare
static int prepare_host(rc_handle_t h)
{
    if (h->host_ready) return RC_OK;
    if (h->kind < 0) return fail(h, RC_ERR_STATE, "no matrix loaded");
    if (h->perm.empty()) {
        h->perm.resize((size_t)h->n_user);
        for (int i = 0; i < h->n_user; ++i) h->perm[(size_t)i] = i;
    }
    if (h->kind == RC_INPUT_LAPLACIAN) {
        permute_sym(h->base, h->perm, h->fac);
        h->sys = h->fac;
        h->trim = false;
    } else if (h->kind == RC_INPUT_GROUNDED) {
        std::vector<int> full(h->perm);
        full.push_back(h->base.n - 1);                 // ground stays last
        permute_sym(h->base, full, h->fac);
        trim_last(h->fac, h->sys);
        h->trim = true;
    } else {
        Csc sp;
        permute_sym(h->base, h->perm, sp);             // order, then append (write_graph.jl:333-383)
        int g = ground_append(sp, h->fac);
        if (g < 0) return fail(h, g, "SDDM input is not diagonally dominant");
        h->trim = (g == 1);
        h->sys = std::move(sp);
    }
    h->nf = h->fac.n;
    h->ns = h->sys.n;
    if (h->sys.nnz() > 2147483647LL || h->fac.nnz() > 2147483647LL)
        return fail(h, RC_ERR_ARG, "nnz exceeds int32 indexing");
    h->host_ready = true;
    return RC_OK;
}

static int prepare(rc_handle_t h)
{
    if (h->prepared) return RC_OK;
    if (h->host_only) return fail(h, RC_ERR_NO_DEVICE, "host-only handle");
    if (h->kind < 0) return fail(h, RC_ERR_STATE, "no matrix loaded");
    const auto t0 = std::chrono::steady_clock::now();
    if (h->perm.empty()) {
        h->perm.resize((size_t)h->n_user);
        for (int i = 0; i < h->n_user; ++i) h->perm[(size_t)i] = i;
    }
    // permuted matrix, ground append / trim, lower slice and dependency
    // counts, all on the device (prep_device.cpp)
    int rc = prepare_device(h);
    if (rc != RC_OK) return rc;
    const int n = h->ns;
    if (h->nsys > 0) {
        // system of every row (user and factored numbering) and the
        // reduction-block table: system s owns q in [sys_off[s], sys_off[s+1])
        // of the system-major position list
        std::vector<int> sys_fac((size_t)n);
        for (int i = 0; i < n; ++i) sys_fac[(size_t)i] = h->sys_user[(size_t)h->perm[(size_t)i]];
        std::vector<int> lo, hi, bs, b0((size_t)h->nsys + 1);
        for (int q = 0; q < h->nsys; ++q) {
            b0[(size_t)q] = (int)lo.size();
            for (int64_t a0 = h->sys_off[(size_t)q]; a0 < h->sys_off[(size_t)q + 1]; a0 += kBatchRows) {
                lo.push_back((int)a0);
                hi.push_back((int)std::min<int64_t>(a0 + kBatchRows, h->sys_off[(size_t)q + 1]));
                bs.push_back(q);
            }
        }
        b0[(size_t)h->nsys] = (int)lo.size();
        h->nblk = (int)lo.size();
        std::vector<int> tab;
        tab.reserve(3 * lo.size() + b0.size());
        tab.insert(tab.end(), lo.begin(), lo.end());
        tab.insert(tab.end(), hi.begin(), hi.end());
        tab.insert(tab.end(), bs.begin(), bs.end());
        tab.insert(tab.end(), b0.begin(), b0.end());
        RC_HIP(h->sys_user_d.ensure((size_t)n * sizeof(int)));
        RC_HIP(h->sys_fac.ensure((size_t)n * sizeof(int)));
        RC_HIP(h->blk_tab.ensure(tab.size() * sizeof(int)));
        RC_HIP(h->bt_state.ensure(batch_state_bytes(h->nsys, h->nblk)));
        RC_HIP(hipMemcpyAsync(h->sys_user_d.p, h->sys_user.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice,
                              h->stream));
        RC_HIP(hipMemcpyAsync(h->sys_fac.p, sys_fac.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, h->stream));
        RC_HIP(hipMemcpyAsync(h->blk_tab.p, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice, h->stream));
        RC_HIP(hipStreamSynchronize(h->stream));
    }
    RC_HIP(hipStreamSynchronize(h->stream));
    h->prepared = true;
    h->factored = false;
    h->rhs_set = false;
    h->prepare_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    invalidate_graph(h);
    return RC_OK;
}

static int rc_get_prepared_csc_impl(rc_handle_t h, int64_t *n, int64_t *nnz, int64_t *colptr, int *rowind, double *val)
{
    RC_CHECK_HANDLE();
    int rc = prepare_host(h);
    if (rc != RC_OK) return rc;
    if (n) *n = h->fac.n;
    if (nnz) *nnz = h->fac.nnz();
    if (colptr) std::memcpy(colptr, h->fac.ptr.data(), h->fac.ptr.size() * sizeof(int64_t));
    if (rowind) std::memcpy(rowind, h->fac.ind.data(), h->fac.ind.size() * sizeof(int));
    if (val) std::memcpy(val, h->fac.val.data(), h->fac.val.size() * sizeof(double));
    return RC_OK;
}

static int rc_write_prepared_mtx_impl(rc_handle_t h, const char *path)
{
    RC_CHECK_HANDLE();
    if (!path) return fail(h, RC_ERR_ARG, "null path");
    int rc = prepare_host(h);
    if (rc != RC_OK) return rc;
    return write_mtx_csc(path, h->fac.n, h->fac.n, h->fac.ptr.data(), h->fac.ind.data(), h->fac.val.data(), h->err);
}

// ------------------------------------------------------------------ factor
static int alloc_factor_work(rc_handle_t h)
{
    const size_t nf = (size_t)h->nf;
    RC_HIP(h->dep.ensure(nf * sizeof(int)));
    RC_HIP(h->head.ensure(nf * kFillLists * sizeof(int)));
    RC_HIP(h->level.ensure(nf * sizeof(int)));
    RC_HIP(h->list0.ensure((size_t)list_capacity(h->nf) * sizeof(int)));
    RC_HIP(h->list1.ensure((size_t)list_capacity(h->nf) * sizeof(int)));
    RC_HIP(h->seg_count.ensure(3 * kListSlots * sizeof(unsigned int)));
    RC_HIP(h->round_count.ensure((nf + 2) * sizeof(unsigned int)));
    RC_HIP(h->big_count.ensure((nf + 2) * sizeof(unsigned int)));
    RC_HIP(h->big_list.ensure(nf * sizeof(int)));
    RC_HIP(h->wave_state.ensure((size_t)(h->num_blocks + kBigGrid) * 4 * sizeof(long long)));
    RC_HIP(h->out_start.ensure(nf * sizeof(long long)));
    RC_HIP(h->out_cnt.ensure(nf * sizeof(int)));
    RC_HIP(h->diag64.ensure(nf * sizeof(double)));
    RC_HIP(h->ctr.ensure(kCtrCount * sizeof(unsigned long long)));
    RC_HIP(h->pool.ensure((size_t)h->pool_cap * sizeof(FillRec)));
    RC_HIP(h->out_row.ensure((size_t)h->out_cap * sizeof(int)));
    RC_HIP(h->out_val.ensure((size_t)h->out_cap * sizeof(double)));
    // only the deferred-column kernel (kBigGrid workgroups) uses the scratch
    RC_HIP(h->scratch.ensure((size_t)kBigGrid * (size_t)h->scratch_cap * (size_t)factor_scratch_bytes_per_entry()));
    RC_HIP(h->aq.ensure(nf * sizeof(int)));
    RC_HIP(h->aq_ctr.ensure(kAqWords * sizeof(unsigned int)));
    return RC_OK;
}

static FactorArgs factor_args(rc_handle_t h)
{
    FactorArgs a;
    a.n = h->nf;
    a.lo_ptr = h->lo_ptr.as<int>();
    a.lo_row = h->lo_row.as<int>();
    a.lo_w = h->lo_w.as<double>();
    a.dep = h->dep.as<int>();
    a.head = h->head.as<int>();
    a.level = h->level.as<int>();
    a.list0 = h->list0.as<int>();
    a.list1 = h->list1.as<int>();
    a.seg_count = h->seg_count.as<unsigned int>();
    a.seg_cap = (int)list_seg_cap(h->nf);
    if (h->seg_cap_limit > 0 && h->seg_cap_limit < a.seg_cap) a.seg_cap = h->seg_cap_limit;   // smaller segments fit the lists
    a.round_count = h->round_count.as<unsigned int>();
    a.wave_state = h->wave_state.as<long long>();
    a.ctr = h->ctr.as<unsigned long long>();
    a.pool = h->pool.as<FillRec>();
    a.pool_cap = h->pool_cap;
    a.out_row = h->out_row.as<int>();
    a.out_val = h->out_val.as<double>();
    a.out_cap = h->out_cap;
    a.out_start = h->out_start.as<long long>();
    a.out_cnt = h->out_cnt.as<int>();
    a.diag = h->diag64.as<double>();
    a.big_list = h->big_list.as<int>();
    a.big_count = h->big_count.as<unsigned int>();
    a.big_ws = h->num_blocks;
    a.scratch = h->scratch.as<char>();
    a.scratch_cap = h->scratch_cap;
    a.seed = h->opt.seed;
    a.aq = h->aq.as<int>();
    a.aq_total = 0;
    a.aq_ctr = h->aq_ctr.as<unsigned int>();
    a.spin_ticks = h->spin_ticks;
    return a;
}

static int run_elimination(rc_handle_t h, double *ms_out, int *regrowths)
{
    const long long slack = factor_pool_slack(h->num_blocks);
    // fill records are addressed by int32 indices (FillRec.next, list heads),
    // so the fill pool holds at most kMaxPoolRecords records
    const long long want =
        std::max<long long>(1024, (long long)std::ceil(h->opt.pool_factor * (double)h->nnz_low)) + slack;
    if (h->pool_cap < want) h->pool_cap = want;
    if (h->pool_cap > kMaxPoolRecords) h->pool_cap = kMaxPoolRecords;
    h->out_cap = std::max(h->out_cap, h->pool_cap + h->nnz_low + slack);
    if (h->scratch_cap < 8192) h->scratch_cap = 8192;
    *regrowths = 0;
    hipStream_t s = h->stream;
    for (int attempt = 0; attempt < 8; ++attempt) {
        int rc = alloc_factor_work(h);
        if (rc != RC_OK) return rc;
        FactorArgs a = factor_args(h);
        // the round and queue kernels read their arguments from a device copy
        RC_HIP(h->fargs.ensure(sizeof(FactorArgs)));
        RC_HIP(hipMemcpyAsync(h->fargs.p, &a, sizeof(FactorArgs), hipMemcpyHostToDevice, s));
        const FactorArgs *da = h->fargs.as<FactorArgs>();
        hipEvent_t e0, e1;
        RC_HIP(hipEventCreate(&e0));
        RC_HIP(hipEventCreate(&e1));
        RC_HIP(hipEventRecord(e0, s));
        RC_HIP(hipMemsetAsync(h->ctr.p, 0, kCtrCount * sizeof(unsigned long long), s));
        RC_HIP(hipMemsetAsync(h->round_count.p, 0, ((size_t)h->nf + 2) * sizeof(unsigned int), s));
        RC_HIP(hipMemsetAsync(h->seg_count.p, 0, 3 * kListSlots * sizeof(unsigned int), s));
        RC_HIP(hipMemsetAsync(h->big_count.p, 0, ((size_t)h->nf + 2) * sizeof(unsigned int), s));
        RC_HIP(hipMemsetAsync(h->wave_state.p, 0, (size_t)(h->num_blocks + kBigGrid) * 4 * sizeof(long long), s));
        launch_factor_setup(a, h->dep_init.as<int>(), s);
        // rounds in growing batches; a round whose list is empty returns at once
        int round = 0, batch = 8;
        unsigned long long st = 0;
        unsigned long long async_total = 0;      // columns eliminated by the queue-driven tail
        int async_round = -1;
        unsigned int *probe = nullptr;     // the next round's list counters, then the status word
        RC_HIP(hipHostMalloc((void **)&probe, (kListSlots + 2) * sizeof(unsigned int), hipHostMallocDefault));
        while (true) {
            int nr = std::min(batch, h->nf + 1 - round);
            if (nr <= 0) break;
            launch_factor_rounds(a, da, h->num_blocks, round, nr, s);
            round += nr;
            hipError_t e = hipMemcpyAsync(probe, h->seg_count.as<unsigned int>() + (size_t)(round % 3) * kListSlots,
                                          kListSlots * sizeof(unsigned int), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipMemcpyAsync(probe + kListSlots, h->ctr.as<unsigned long long>() + kCtrStatus,
                                                    sizeof(unsigned long long), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) { (void)hipHostFree(probe); return fail(h, RC_ERR_HIP, hipGetErrorString(e)); }
            std::memcpy(&st, probe + kListSlots, sizeof(st));
            unsigned long long pending = probe[kListSegs];
            for (int q = 0; q < kListSegs; ++q) pending += std::min<unsigned long long>(probe[q], (unsigned long long)a.seg_cap);
            if (st != 0 || pending == 0) break;
            if (h->async_tail && pending <= (unsigned long long)h->async_width) {
                // narrow from here: the rest runs in one queue-driven launch
                std::vector<unsigned int> cnt((size_t)round);
                e = hipMemcpyAsync(cnt.data(), h->round_count.p, cnt.size() * sizeof(unsigned int),
                                   hipMemcpyDeviceToHost, s);
                if (e == hipSuccess) e = hipStreamSynchronize(s);
                if (e != hipSuccess) { (void)hipHostFree(probe); return fail(h, RC_ERR_HIP, hipGetErrorString(e)); }
                unsigned long long before = 0;
                for (unsigned int v : cnt) before += v;
                async_total = (unsigned long long)(h->nf - 1) - before;
                async_round = round;
                const unsigned int aqt = (unsigned int)async_total;
                e = hipMemcpyAsync((char *)h->fargs.p + offsetof(FactorArgs, aq_total), &aqt, sizeof(aqt),
                                   hipMemcpyHostToDevice, s);
                if (e != hipSuccess) { (void)hipHostFree(probe); return fail(h, RC_ERR_HIP, hipGetErrorString(e)); }
                launch_factor_async(a, da, round, aqt, s);
                break;
            }
            batch = std::min(batch * 2, 64);     // bounds the empty rounds launched past the end
        }
        (void)hipHostFree(probe);
        launch_factor_finish(a, s);
        RC_HIP(hipGetLastError());
        RC_HIP(hipEventRecord(e1, s));
        unsigned long long c[kCtrCount];
        RC_HIP(hipMemcpyAsync(c, h->ctr.p, sizeof(c), hipMemcpyDeviceToHost, s));
        std::vector<unsigned int> counts((size_t)round + 1);
        RC_HIP(hipMemcpyAsync(counts.data(), h->round_count.p, counts.size() * sizeof(unsigned int),
                              hipMemcpyDeviceToHost, s));
        unsigned int aq_tail = 0;
        if (async_round >= 0)
            RC_HIP(hipMemcpyAsync(&aq_tail, h->aq_ctr.as<unsigned int>() + kAqTail, sizeof(aq_tail),
                                  hipMemcpyDeviceToHost, s));
        RC_HIP(hipStreamSynchronize(s));
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms_out = ms;
        st = c[kCtrStatus];
        if (st == 0) {
            unsigned long long done = async_total;
            int used = 0;
            const size_t nc = async_round >= 0 ? (size_t)async_round : counts.size();
            for (size_t i = 0; i < nc; ++i) {
                done += counts[i];
                if (counts[i]) used = (int)i + 1;
            }
            // the queue-driven tail queued every column it was given
            if (async_round >= 0 && aq_tail != async_total)
                return fail(h, RC_ERR_STATE, "queue-driven elimination queued " + std::to_string(aq_tail) + " of " +
                                                 std::to_string(async_total) + " columns");
            if (done != (unsigned long long)(h->nf - 1))
                return fail(h, RC_ERR_STATE, "elimination did not retire every column (" + std::to_string(done) + "/" +
                                                 std::to_string(h->nf - 1) + ")");
            h->rounds = used;
            if (c[kCtrNnzOff] > (unsigned long long)kMaxFactorEntries)
                return fail(h, RC_ERR_POOL, "factor has more than 2^31 - 2 off-diagonal entries (int32 CSR)");
            h->fstats.nnz_factor = (int64_t)c[kCtrNnzOff] + h->nf;
            h->fstats.fills = (int64_t)c[kCtrFills];
            h->fstats.max_degree = (int64_t)c[kCtrMaxDeg];
            h->fstats.zero_pivots = (int32_t)c[kCtrZeroPivots];
            return RC_OK;
        }
        if (st & kStatusInconsistent)
            return fail(h, RC_ERR_STATE, "elimination kernel found an inconsistent fill list");
        if (st & kStatusTimeout) return fail(h, RC_ERR_TIMEOUT, "elimination queue wait timed out");
        if (st & kStatusFillPool) {
            if (h->pool_cap >= kMaxPoolRecords)
                return fail(h, RC_ERR_POOL, "fill pool would exceed 2^31 - 1 records (int32 record indices)");
            h->pool_cap = std::min<long long>(2 * (h->pool_cap - slack) + slack, kMaxPoolRecords);
        }
        if (st & (kStatusOutPool | kStatusFillPool)) h->out_cap = h->pool_cap + h->nnz_low + slack;
        if (st & kStatusScratch) h->scratch_cap = (int)std::min<int64_t>(pow2_at_least((int64_t)c[kCtrScratchNeed]), 1 << 28);
        ++*regrowths;
    }
    return fail(h, RC_ERR_POOL, "elimination pools exhausted after regrowth");
}

}  // extern "C"

static int bit_width(int64_t v)
{
    int b = 0;
    while (v > 0) { ++b; v >>= 1; }
    return b;
}

template <typename T>
static int assemble(rc_handle_t h, double *ms_out)
{
    h->multi_ready = false;
    hipStream_t s = h->stream;
    hipEvent_t e0, e1;
    RC_HIP(hipEventCreate(&e0));
    RC_HIP(hipEventCreate(&e1));
    RC_HIP(hipEventRecord(e0, s));
    const int ns = h->ns;
    AssembleArgs a;
    a.n = h->nf;
    a.ns = ns;
    a.trim = h->trim ? 1 : 0;
    a.out_start = h->out_start.as<long long>();
    a.out_cnt = h->out_cnt.as<int>();
    a.out_row = h->out_row.as<int>();
    a.out_val = h->out_val.as<double>();
    a.diag_in = h->diag64.as<double>();
    a.level = h->level.as<int>();
    const size_t nsz = (size_t)ns + 1;

    // Work buffers sized once for the whole assembly (levels <= ns, chunks <= ns,
    // tiles <= ns/64 + 1), so nothing is reallocated while kernels that read
    // it may still be in flight.
    RC_HIP(h->tmp_a.ensure(3 * (nsz + 1) * sizeof(int)));
    RC_HIP(h->tmp_scan.ensure(scan_tmp_elems(ns + 2) * sizeof(int) + 64));

    // 1. CSR of L^T in the factored numbering (ground row dropped)
    RC_HIP(h->lt_ptr.ensure(nsz * sizeof(int)));
    launch_lt_counts(a, h->tmp_a.as<int>(), s);
    scan_exclusive(h->tmp_a.as<int>(), h->lt_ptr.as<int>(), ns, h->tmp_scan.as<int>(), s);
    int nnz = 0;
    RC_HIP(hipMemcpyAsync(&nnz, h->lt_ptr.as<int>() + ns, sizeof(int), hipMemcpyDeviceToHost, s));
    RC_HIP(hipStreamSynchronize(s));
    h->nnz_lt = nnz;
    RC_HIP(h->lt_ind.ensure((size_t)std::max(nnz, 1) * sizeof(int)));
    RC_HIP(h->lt_val.ensure((size_t)std::max(nnz, 1) * sizeof(T)));
    RC_HIP(h->diagT.ensure((size_t)std::max(ns, 1) * sizeof(T)));
    launch_lt_fill<T>(a, h->lt_ptr.as<int>(), h->lt_ind.as<int>(), h->lt_val.as<T>(), h->diagT.as<T>(), s);

    // 2. rows of L: stable radix sort of the entries by row (columns stay ascending)
    const size_t big = (size_t)std::max(nnz, ns) + 1;
    RC_HIP(h->keys0.ensure(big * sizeof(uint32_t)));
    RC_HIP(h->keys1.ensure(big * sizeof(uint32_t)));
    RC_HIP(h->vals0.ensure(big * sizeof(int)));
    RC_HIP(h->vals1.ensure(big * sizeof(int)));
    RC_HIP(h->rowk.ensure(big * sizeof(int)));
    RC_HIP(h->l_ptr.ensure(nsz * sizeof(int)));
    const size_t sort_tmp = radix_tmp_ints((int)big);
    RC_HIP(h->tmp_b.ensure(std::max(sort_tmp, nsz * 2) * sizeof(int)));
    uint32_t *k0 = h->keys0.as<uint32_t>(), *k1 = h->keys1.as<uint32_t>();
    launch_copy_keys(h->lt_ind.as<int>(), k0, nnz, s);
    launch_iota(h->vals0.as<int>(), nnz, s);
    radix_sort_pairs(k0, h->vals0.as<int>(), k1, h->vals1.as<int>(), nnz, bit_width(ns), h->tmp_b.as<int>(), s);
    launch_segment_bounds(nnz, k0, 0, ns, h->l_ptr.as<int>(), s);

    // 3. level order: rows sorted by (level, descending length of their L row)
    unsigned long long *d_max = h->ctr.as<unsigned long long>() + kCtrCount - 1;
    RC_HIP(hipMemsetAsync(d_max, 0, sizeof(unsigned long long), s));
    launch_level_max(ns, h->level.as<int>(), d_max, s);
    unsigned long long maxlev = 0;
    RC_HIP(hipMemcpyAsync(&maxlev, d_max, sizeof(maxlev), hipMemcpyDeviceToHost, s));
    RC_HIP(hipStreamSynchronize(s));
    const int nl = ns > 0 ? (int)maxlev + 1 : 0;
    h->nlevels = nl;
    const int lvl_bits = bit_width(nl);
    // batched systems: (level, system, length) keys keep a system's rows of a
    // level contiguous, so the segmented reductions read runs of positions
    const int sys_bits = (h->nsys > 1 && lvl_bits + bit_width(h->nsys - 1) <= 32) ? bit_width(h->nsys - 1) : 0;
    int len_bits = std::max(0, std::min(16, 32 - lvl_bits - sys_bits));
    RC_HIP(h->order.ensure(nsz * sizeof(int)));
    RC_HIP(h->pos_of.ensure(nsz * sizeof(int)));
    // rows of a level: in the input order (spatial) or, under the bisection
    // locality order and for systems of more than 4 M rows, in width classes
    // each in that order (measured: 256^3 sweeps 2.93 / 2.59 -> 2.80 / 2.52 ms;
    // at 128^3 the classes cost more than they gain, 0.77 -> 0.84 ms forward)
    const bool classes_by_rank = h->level_order == 1 && !h->loc_rank.empty() && !h->level_order_set &&
                                 h->nsys == 0 && ns > (1 << 22);
    const int level_order = classes_by_rank ? 3 : h->level_order;
    if (h->nsys > 0) {
        launch_level_keys_sys(ns, h->level.as<int>(), h->l_ptr.as<int>(), len_bits, h->sys_fac.as<int>(), sys_bits, k0,
                              s);
        launch_iota(h->order.as<int>(), ns, s);
        radix_sort_pairs(k0, h->order.as<int>(), k1, h->pos_of.as<int>(), ns, lvl_bits + sys_bits + len_bits,
                         h->tmp_b.as<int>(), s);
    } else if (level_order == 3 && lvl_bits + 2 <= 32) {
        // rows of a level in four classes by the lanes per row their width
        // needs, each class in the loaded (user) order: short rows get their
        // own (wider) chunks instead of the chunk size of the level's widest
        // row, and neighbouring rows of a chunk stay neighbours in the input
        if (!h->loc_rank.empty()) {      // order = rows by locality rank
            launch_locality_keys(ns, h->perm_d.as<int>(), h->n_user, h->loc_d.as<int>(), k0, s);
            launch_iota(h->order.as<int>(), ns, s);
            radix_sort_pairs(k0, h->order.as<int>(), k1, h->pos_of.as<int>(), ns, bit_width(h->n_user),
                             h->tmp_b.as<int>(), s);
        } else {
            launch_inverse_perm(ns, h->perm_d.as<int>(), h->order.as<int>(), s);    // order = rows by user index
        }
        launch_level_group_keys(ns, h->order.as<int>(), h->level.as<int>(), h->l_ptr.as<int>(), h->lt_ptr.as<int>(),
                                k0, s);
        radix_sort_pairs(k0, h->order.as<int>(), k1, h->pos_of.as<int>(), ns, lvl_bits + 2, h->tmp_b.as<int>(), s);
        len_bits = 2;
    } else if (level_order == 2 && lvl_bits + 3 <= 32) {
        // rows of a level in classes of similar width (log2 of the wider of
        // the L and L^T rows, widest first), each class in the loaded (user)
        // order: chunks stay near-uniform in width and neighbouring rows of a
        // chunk are neighbours in the input numbering
        launch_inverse_perm(ns, h->perm_d.as<int>(), h->order.as<int>(), s);    // order = rows by user index
        launch_level_bucket_keys(ns, h->order.as<int>(), h->level.as<int>(), h->l_ptr.as<int>(),
                                 h->lt_ptr.as<int>(), k0, s);
        radix_sort_pairs(k0, h->order.as<int>(), k1, h->pos_of.as<int>(), ns, lvl_bits + 3, h->tmp_b.as<int>(), s);
        len_bits = 3;
    } else if (level_order == 1) {
        // rows of a level kept in the loaded (user) order: neighbouring rows of
        // a chunk are neighbours in the input numbering, so gathers share lines
        if (!h->loc_rank.empty()) {
            // rows in the recursive bisection order of the input graph
            launch_locality_keys(ns, h->perm_d.as<int>(), h->n_user, h->loc_d.as<int>(), k0, s);
            launch_iota(h->order.as<int>(), ns, s);
            radix_sort_pairs(k0, h->order.as<int>(), k1, h->pos_of.as<int>(), ns, bit_width(h->n_user),
                             h->tmp_b.as<int>(), s);
        } else {
            launch_inverse_perm(ns, h->perm_d.as<int>(), h->order.as<int>(), s);    // order = rows by user index
        }
        launch_gather_i(h->level.as<int>(), h->order.as<int>(), (int *)k0, ns, s);
        radix_sort_pairs(k0, h->order.as<int>(), k1, h->pos_of.as<int>(), ns, lvl_bits, h->tmp_b.as<int>(), s);
        launch_level_keys(ns, h->level.as<int>(), h->l_ptr.as<int>(), len_bits, k1, s);
        launch_gather_i((const int *)k1, h->order.as<int>(), (int *)k0, ns, s);   // sorted keys for the bounds
    } else {
        launch_level_keys(ns, h->level.as<int>(), h->l_ptr.as<int>(), len_bits, k0, s);
        launch_iota(h->order.as<int>(), ns, s);
        radix_sort_pairs(k0, h->order.as<int>(), k1, h->pos_of.as<int>(), ns, lvl_bits + len_bits, h->tmp_b.as<int>(),
                         s);
    }
    launch_inverse_perm(ns, h->order.as<int>(), h->pos_of.as<int>(), s);
    // position of the L^T row of every entry (the column of its L row): the
    // forward layouts read it directly instead of mapping per entry
    launch_row_expand(ns, h->lt_ptr.as<int>(), h->pos_of.as<int>(), h->rowk.as<int>(), s);
    const size_t nlsz = (size_t)nl + 1;
    // chunk segments: the levels, or, in the width-class order, every
    // (level, class) run of a level (8 per level), so each class gets the
    // chunk size of its own widest row rather than of the level's
    const bool classes = sys_bits == 0 && ((level_order == 2 && len_bits == 3) || (level_order == 3 && len_bits == 2));
    const int per_level = classes ? 1 << len_bits : 1;
    const int nseg = nl * per_level;
    const int seg_shift = classes ? 0 : len_bits + sys_bits;
    const size_t nssz = (size_t)nseg + 1;
    RC_HIP(h->seg_work.ensure(3 * nssz * sizeof(int)));
    int *lptr = h->seg_work.as<int>();
    int *lcnt = lptr + nssz;
    int *cptr = lcnt + nssz;
    launch_segment_bounds(ns, k0, seg_shift, nseg, lptr, s);
    // chunk size per segment from its widest row (at most 8 entries per lane)
    RC_HIP(h->lvl_w.ensure(nssz * sizeof(int)));
    RC_HIP(hipMemsetAsync(h->lvl_w.p, 0, nssz * sizeof(int), s));
    launch_level_width(ns, h->order.as<int>(), h->level.as<int>(), classes ? k0 : nullptr, seg_shift,
                       h->l_ptr.as<int>(), h->lt_ptr.as<int>(), h->lvl_w.as<int>(), s);
    RC_HIP(h->tmp_scan.ensure(scan_tmp_elems(nseg + 2) * sizeof(int) + 64));
    launch_chunk_count(nseg, lptr, h->lvl_w.as<int>(), h->sweep_gmin, lcnt, s);
    scan_exclusive(lcnt, cptr, nseg, h->tmp_scan.as<int>(), s);
    int nchunks = 0;
    RC_HIP(hipMemcpyAsync(&nchunks, cptr + nseg, sizeof(int), hipMemcpyDeviceToHost, s));
    RC_HIP(hipStreamSynchronize(s));
    h->nchunks = nchunks;
    RC_HIP(h->chunk_start.ensure((size_t)std::max(nchunks, 1) * sizeof(int)));
    RC_HIP(h->chunk_cnt.ensure((size_t)std::max(nchunks, 1) * sizeof(int)));
    launch_chunk_fill(nchunks, nseg, lptr, cptr, h->lvl_w.as<int>(), h->sweep_gmin, h->chunk_start.as<int>(),
                      h->chunk_cnt.as<int>(), s);
    // level split: the trailing run of levels with fewer than tail_max_rows
    // rows, as long as their rows fit the tail workgroup's LDS, is the tail
    std::vector<int> hl(nlsz), hc(nlsz);
    {
        std::vector<int> hs(nssz), hcs(nssz);
        RC_HIP(hipMemcpyAsync(hs.data(), lptr, nssz * sizeof(int), hipMemcpyDeviceToHost, s));
        RC_HIP(hipMemcpyAsync(hcs.data(), cptr, nssz * sizeof(int), hipMemcpyDeviceToHost, s));
        RC_HIP(hipStreamSynchronize(s));
        for (size_t L = 0; L < nlsz; ++L) {       // level L starts at its first segment
            hl[L] = hs[L * (size_t)per_level];
            hc[L] = hcs[L * (size_t)per_level];
        }
        if (h->tail_lds_cap < 0) h->tail_lds_cap = tail_lds_limit(h->opt.device);
        int tl = nl;
        int64_t rows = 0;
        while (tl > 0) {
            const int c = hl[(size_t)tl] - hl[(size_t)tl - 1];
            // values + one descriptor per tail level + the empty descriptor
            const int64_t bytes = ((rows + c + 1) * (int64_t)sizeof(T) + 15) / 16 * 16 + 16 * (int64_t)(nl - tl + 2);
            if (c >= h->tail_max_rows || bytes > (int64_t)h->tail_lds_cap) break;
            rows += c;
            --tl;
        }
        h->tail_level = tl;
        h->bulk_chunks = nl > 0 ? hc[(size_t)tl] : 0;
    }

    // 4. sliced ELL of L (forward) and L^T (backward) over the chunks
    const size_t csz = (size_t)nchunks + 1;
    RC_HIP(h->lf_width.ensure(csz * sizeof(int)));
    RC_HIP(h->lf_off.ensure(csz * sizeof(int)));
    RC_HIP(h->lb_width.ensure(csz * sizeof(int)));
    RC_HIP(h->lb_off.ensure(csz * sizeof(int)));
    auto ell_offsets = [&](const int *width, int *off, int nslices, int64_t &total) -> int {
        launch_mul64(width, h->tmp_a.as<int>(), nslices, s);
        scan_exclusive(h->tmp_a.as<int>(), off, nslices, h->tmp_scan.as<int>(), s);
        int t = 0;
        RC_HIP(hipMemcpyAsync(&t, off + nslices, sizeof(int), hipMemcpyDeviceToHost, s));
        RC_HIP(hipStreamSynchronize(s));
        total = t;
        return RC_OK;
    };
    // grouped layout: width -> slice count in place, slice block sizes -> offsets
    auto sweep_offsets = [&](const int *ptr, int *slices, int *off, int64_t &total) -> int {
        launch_ell_width(nchunks, h->chunk_start.as<int>(), h->chunk_cnt.as<int>(), h->order.as<int>(), ns, ptr,
                         slices, s);
        launch_sweep_layout(nchunks, h->chunk_cnt.as<int>(), slices, slices, h->tmp_a.as<int>(), s);
        scan_exclusive(h->tmp_a.as<int>(), off, nchunks, h->tmp_scan.as<int>(), s);
        int t = 0;
        RC_HIP(hipMemcpyAsync(&t, off + nchunks, sizeof(int), hipMemcpyDeviceToHost, s));
        RC_HIP(hipStreamSynchronize(s));
        total = t;
        return RC_OK;
    };
    int rc = sweep_offsets(h->l_ptr.as<int>(), h->lf_width.as<int>(), h->lf_off.as<int>(), h->ell_l);
    if (rc != RC_OK) return rc;
    rc = sweep_offsets(h->lt_ptr.as<int>(), h->lb_width.as<int>(), h->lb_off.as<int>(), h->ell_u);
    if (rc != RC_OK) return rc;
    // one more entry after the chunks: the zero slot (index ns, value 0)
    // that lanes without a slot load when the sweeps prefetch a chunk
    RC_HIP(h->lf_ind.ensure(((size_t)h->ell_l + 1) * sizeof(int)));
    RC_HIP(h->lf_val.ensure(((size_t)h->ell_l + 1) * sizeof(T)));
    RC_HIP(h->lb_ind.ensure(((size_t)h->ell_u + 1) * sizeof(int)));
    RC_HIP(h->lb_val.ensure(((size_t)h->ell_u + 1) * sizeof(T)));
    {
        const T zero_val = (T)0;
        RC_HIP(hipMemcpyAsync(h->lf_ind.as<int>() + h->ell_l, &ns, sizeof(int), hipMemcpyHostToDevice, s));
        RC_HIP(hipMemcpyAsync(h->lf_val.as<T>() + h->ell_l, &zero_val, sizeof(T), hipMemcpyHostToDevice, s));
        RC_HIP(hipMemcpyAsync(h->lb_ind.as<int>() + h->ell_u, &ns, sizeof(int), hipMemcpyHostToDevice, s));
        RC_HIP(hipMemcpyAsync(h->lb_val.as<T>() + h->ell_u, &zero_val, sizeof(T), hipMemcpyHostToDevice, s));
        RC_HIP(hipStreamSynchronize(s));
    }
    EllSrc fs{h->l_ptr.as<int>(), nullptr, h->lt_val.p, h->vals0.as<int>(), h->rowk.as<int>(), nullptr};
    launch_sweep_fill<T, T>(nchunks, h->chunk_start.as<int>(), h->chunk_cnt.as<int>(), h->order.as<int>(), fs,
                            h->lf_off.as<int>(), h->lf_width.as<int>(), ns, (T)0, h->lf_ind.as<int>(), h->lf_val.as<T>(),
                            s);
    EllSrc bs{h->lt_ptr.as<int>(), h->lt_ind.as<int>(), h->lt_val.p, nullptr, nullptr, h->pos_of.as<int>()};
    launch_sweep_fill<T, T>(nchunks, h->chunk_start.as<int>(), h->chunk_cnt.as<int>(), h->order.as<int>(), bs,
                            h->lb_off.as<int>(), h->lb_width.as<int>(), ns, (T)0, h->lb_ind.as<int>(), h->lb_val.as<T>(),
                            s);
    RC_HIP(h->diagP.ensure((size_t)std::max(ns, 1) * sizeof(T)));
    launch_gather_t<T>(h->diagT.as<T>(), h->order.as<int>(), h->diagP.as<T>(), ns, s);
    RC_HIP(h->cm_f.ensure((size_t)std::max(nchunks, 1) * sizeof(int4)));
    RC_HIP(h->cm_b.ensure((size_t)std::max(nchunks, 1) * sizeof(int4)));
    launch_pack_meta(nchunks, h->chunk_start.as<int>(), h->chunk_cnt.as<int>(), h->lf_off.as<int>(),
                     h->lf_width.as<int>(), h->cm_f.as<int4>(), s);
    launch_pack_meta(nchunks, h->chunk_start.as<int>(), h->chunk_cnt.as<int>(), h->lb_off.as<int>(),
                     h->lb_width.as<int>(), h->cm_b.as<int4>(), s);

    // 4b. the tail: level table, forward bulk-part CSR, internal ELL per direction
    h->ntail_levels = nl - h->tail_level;
    h->tail_pos0 = nl > 0 ? hl[(size_t)h->tail_level] : ns;
    h->ntail = ns - h->tail_pos0;
    h->tail_lds_bytes = tail_value_bytes<T>(h->ntail) + 16 * (h->ntail_levels + 1);
    if (h->ntail > 0) {
        const int ntl = h->ntail_levels, tp0 = h->tail_pos0, nt = h->ntail;
        std::vector<int> ls((size_t)ntl), lc((size_t)ntl);
        for (int L = 0; L < ntl; ++L) {
            ls[(size_t)L] = hl[(size_t)(h->tail_level + L)];
            lc[(size_t)L] = hl[(size_t)(h->tail_level + L + 1)] - ls[(size_t)L];
        }
        RC_HIP(h->tl_start.ensure((size_t)ntl * sizeof(int)));
        RC_HIP(h->tl_cnt.ensure((size_t)ntl * sizeof(int)));
        RC_HIP(hipMemcpyAsync(h->tl_start.p, ls.data(), (size_t)ntl * sizeof(int), hipMemcpyHostToDevice, s));
        RC_HIP(hipMemcpyAsync(h->tl_cnt.p, lc.data(), (size_t)ntl * sizeof(int), hipMemcpyHostToDevice, s));
        RC_HIP(h->tacc.ensure((size_t)nt * sizeof(T)));
        RC_HIP(h->ti_ptr.ensure(((size_t)nt + 1) * sizeof(int)));
        int *cnt_b = h->tmp_a.as<int>(), *cnt_i = cnt_b + nt + 1;
        auto build = [&](const EllSrc &src, bool fwd, DevVec &off, DevVec &slices, DevVec &ind, DevVec &val) -> int {
            launch_tail_count(nt, tp0, h->order.as<int>(), src, cnt_b, cnt_i, s);
            if (fwd) RC_HIP(h->tp_ptr.ensure(((size_t)nt + 1) * sizeof(int)));
            int nb = 0, ni = 0;
            if (fwd) {
                scan_exclusive(cnt_b, h->tp_ptr.as<int>(), nt, h->tmp_scan.as<int>(), s);
                RC_HIP(hipMemcpyAsync(&nb, h->tp_ptr.as<int>() + nt, sizeof(int), hipMemcpyDeviceToHost, s));
            }
            scan_exclusive(cnt_i, h->ti_ptr.as<int>(), nt, h->tmp_scan.as<int>(), s);
            RC_HIP(hipMemcpyAsync(&ni, h->ti_ptr.as<int>() + nt, sizeof(int), hipMemcpyDeviceToHost, s));
            RC_HIP(hipStreamSynchronize(s));
            if (!fwd) {
                // a backward dependency always lies in a deeper level, so the
                // backward tail has no bulk part
                int64_t tot = 0;
                std::vector<int> hb((size_t)nt);
                RC_HIP(hipMemcpy(hb.data(), cnt_b, (size_t)nt * sizeof(int), hipMemcpyDeviceToHost));
                for (int v : hb) tot += v;
                if (tot != 0) return fail(h, RC_ERR_STATE, "backward tail row depends on a wider level");
            }
            if (fwd) {
                RC_HIP(h->tp_ind.ensure((size_t)std::max(nb, 1) * sizeof(int)));
                RC_HIP(h->tp_val.ensure((size_t)std::max(nb, 1) * sizeof(T)));
            }
            RC_HIP(h->ti_ind.ensure((size_t)std::max(ni, 1) * sizeof(int)));
            RC_HIP(h->ti_val.ensure((size_t)std::max(ni, 1) * sizeof(T)));
            launch_tail_split<T>(nt, tp0, h->order.as<int>(), src, fwd ? h->tp_ptr.as<int>() : cnt_b,
                                 h->ti_ptr.as<int>(), fwd ? h->tp_ind.as<int>() : nullptr,
                                 fwd ? h->tp_val.as<T>() : nullptr, h->ti_ind.as<int>(), h->ti_val.as<T>(), s);
            RC_HIP(slices.ensure((size_t)ntl * sizeof(int)));
            RC_HIP(off.ensure(((size_t)ntl + 1) * sizeof(int)));
            int *size = cnt_b;        // reuse: ntl <= nt
            launch_tail_layout(ntl, h->tl_start.as<int>(), h->tl_cnt.as<int>(), tp0, h->ti_ptr.as<int>(),
                               slices.as<int>(), size, s);
            scan_exclusive(size, off.as<int>(), ntl, h->tmp_scan.as<int>(), s);
            int total = 0;
            RC_HIP(hipMemcpyAsync(&total, off.as<int>() + ntl, sizeof(int), hipMemcpyDeviceToHost, s));
            RC_HIP(hipStreamSynchronize(s));
            // one more block after the levels: the padding lanes without a slot load
            RC_HIP(ind.ensure(((size_t)total + kTailBlock) * sizeof(int)));
            RC_HIP(val.ensure(((size_t)total + kTailBlock) * sizeof(T)));
            launch_tail_fill<T>(ntl, h->tl_start.as<int>(), h->tl_cnt.as<int>(), tp0, h->ti_ptr.as<int>(),
                                h->ti_ind.as<int>(), h->ti_val.as<T>(), off.as<int>(), slices.as<int>(), nt,
                                ind.as<int>(), val.as<T>(), s);
            int pad_ind[kTailBlock];
            T pad_val[kTailBlock];
            for (int u = 0; u < kTailBlock; ++u) {
                pad_ind[u] = nt;
                pad_val[u] = (T)0;
            }
            RC_HIP(hipMemcpyAsync(ind.as<int>() + total, pad_ind, sizeof(pad_ind), hipMemcpyHostToDevice, s));
            RC_HIP(hipMemcpyAsync(val.as<T>() + total, pad_val, sizeof(pad_val), hipMemcpyHostToDevice, s));
            RC_HIP(hipStreamSynchronize(s));
            return RC_OK;
        };
        rc = build(fs, true, h->tf_off, h->tf_slices, h->tf_ind, h->tf_val);
        if (rc != RC_OK) return rc;
        rc = build(bs, false, h->tb_off, h->tb_slices, h->tb_ind, h->tb_val);
        if (rc != RC_OK) return rc;
    }

    // 5. A in position numbering, sliced ELL over 64-position tiles
    const int ntiles = (ns + 63) / 64;
    h->ntiles = ntiles;
    const size_t tsz = (size_t)ntiles + 1;
    RC_HIP(h->a_width.ensure(tsz * sizeof(int)));
    RC_HIP(h->a_off.ensure(tsz * sizeof(int)));
    launch_ell_width(ntiles, nullptr, nullptr, h->order.as<int>(), ns, h->a_ptr.as<int>(), h->a_width.as<int>(), s);
    rc = ell_offsets(h->a_width.as<int>(), h->a_off.as<int>(), ntiles, h->ell_a);
    if (rc != RC_OK) return rc;
    RC_HIP(h->a_ell_ind.ensure((size_t)std::max<int64_t>(h->ell_a, 1) * sizeof(int)));
    RC_HIP(h->a_ell_val.ensure((size_t)std::max<int64_t>(h->ell_a, 1) * sizeof(double)));
    EllSrc as{h->a_ptr.as<int>(), h->a_ind.as<int>(), h->a_val.p, nullptr, nullptr, h->pos_of.as<int>()};
    launch_ell_fill<double, double>(ntiles, nullptr, nullptr, h->order.as<int>(), ns, as, h->a_off.as<int>(),
                                    h->a_width.as<int>(), ns, -0.0, h->a_ell_ind.as<int>(), h->a_ell_val.as<double>(), s);
    // user index of every position (RHS gather / solution scatter)
    RC_HIP(h->idx_user.ensure(nsz * sizeof(int)));
    launch_gather_i(h->perm_d.as<int>(), h->order.as<int>(), h->idx_user.as<int>(), ns, s);
    // the SpMV visits the tiles in the user order of their first rows: a row's
    // neighbours sit in other levels, i.e. far apart in positions, but near it
    // in the input numbering, so tiles visited together gather from the same
    // lines (grids: a few neighbouring planes; batches: one system at a time).
    // With the bisection order inside levels a tile is a compact region, and
    // the tiles are visited in the bisection order of their first rows.
    RC_HIP(h->tile_ord.ensure(tsz * sizeof(int)));
    RC_HIP(h->tile_k.ensure(2 * tsz * sizeof(uint32_t)));
    RC_HIP(h->tile_v.ensure(tsz * sizeof(int)));
    RC_HIP(h->tile_tmp.ensure(radix_tmp_ints((int)tsz) * sizeof(int)));
    const int *tile_src = h->idx_user.as<int>();
    if (!h->loc_rank.empty()) {
        RC_HIP(h->loc_pos.ensure(nsz * sizeof(int)));
        launch_locality_keys(ns, h->idx_user.as<int>(), h->n_user, h->loc_d.as<int>(), h->loc_pos.as<uint32_t>(), s);
        tile_src = h->loc_pos.as<int>();
    }
    launch_tile_keys(ntiles, tile_src, h->tile_k.as<uint32_t>(), h->tile_ord.as<int>(), s);
    radix_sort_pairs(h->tile_k.as<uint32_t>(), h->tile_ord.as<int>(), h->tile_k.as<uint32_t>() + tsz,
                     h->tile_v.as<int>(), ntiles, std::max(1, bit_width(ns)), h->tile_tmp.as<int>(), s);
    if (h->nsys > 0) {
        // system of every position, and the positions listed system by system
        RC_HIP(h->sys_pos.ensure(nsz * sizeof(int)));
        RC_HIP(h->spos.ensure(nsz * sizeof(int)));
        launch_gather_i(h->sys_user_d.as<int>(), h->idx_user.as<int>(), h->sys_pos.as<int>(), ns, s);
        launch_copy_keys(h->sys_pos.as<int>(), k0, ns, s);
        launch_iota(h->vals0.as<int>(), ns, s);
        radix_sort_pairs(k0, h->vals0.as<int>(), k1, h->vals1.as<int>(), ns, std::max(1, bit_width(h->nsys - 1)),
                         h->tmp_b.as<int>(), s);
        RC_HIP(hipMemcpyAsync(h->spos.p, h->vals0.p, (size_t)ns * sizeof(int), hipMemcpyDeviceToDevice, s));
    }

    RC_HIP(hipGetLastError());
    RC_HIP(hipEventRecord(e1, s));
    RC_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_out = ms;
    return RC_OK;
}

extern "C" {

static double factor_bytes(rc_handle_t h)
{
    // 12 nnz_low + 4 (n+1) + 2 (4+s) F + (4+s) nnz_off + s n, s = 8 (elimination in FP64)
    const double s = 8.0;
    double F = (double)h->fstats.fills;
    double off = (double)(h->fstats.nnz_factor - h->nf);
    return 12.0 * (double)h->nnz_low + 4.0 * (double)(h->nf + 1) + 2.0 * (4.0 + s) * F + (4.0 + s) * off +
           s * (double)h->nf;
}

static int rc_factor_impl(rc_handle_t h, rc_factor_stats *st)
{
    RC_CHECK_HANDLE();
    if (h->host_only) return fail(h, RC_ERR_NO_DEVICE, "host-only handle");
    RC_HIP(hipSetDevice(h->opt.device));
    const bool had = h->prepared;
    int rc = prepare(h);
    if (rc != RC_OK) return rc;
    const double prep_ms = had ? 0.0 : h->prepare_ms;
    std::memset(&h->fstats, 0, sizeof(h->fstats));
    double fms = 0.0, ams = 0.0;
    int regrow = 0;
    h->factored = false;
    h->m_solved = false;
    h->m_ran = false;
    invalidate_graph(h);
    rc = run_elimination(h, &fms, &regrow);
    if (rc != RC_OK) return rc;
    h->prec = h->opt.factor_precision;
    rc = (h->prec == 32) ? assemble<float>(h, &ams) : assemble<double>(h, &ams);
    if (rc != RC_OK) return rc;
    h->fstats.n = h->nf;
    h->fstats.nnz_lower = h->nnz_low;
    h->fstats.nlevels = h->nlevels;
    h->fstats.factor_ms = fms;
    h->fstats.assemble_ms = ams;
    h->fstats.pool_regrowths = regrow;
    h->fstats.bytes_factor = factor_bytes(h);
    h->fstats.prepare_ms = prep_ms;
    h->factored = true;
    if (st) *st = h->fstats;
    return RC_OK;
}

// ------------------------------------------------------------------ IC(0) factor
// The context preconditioner: incomplete Cholesky IC(0) of the prepared
// matrix (ic0.hip) written into the elimination's output layout, then the
// same assembly, so every solve entry point runs with it unchanged.
static int rc_factor_ic0_impl(rc_handle_t h, rc_factor_stats *st)
{
    RC_CHECK_HANDLE();
    if (h->host_only) return fail(h, RC_ERR_NO_DEVICE, "host-only handle");
    RC_HIP(hipSetDevice(h->opt.device));
    const bool had = h->prepared;
    int rc = prepare(h);
    if (rc != RC_OK) return rc;
    const double prep_ms = had ? 0.0 : h->prepare_ms;
    std::memset(&h->fstats, 0, sizeof(h->fstats));
    h->factored = false;
    h->m_solved = false;
    h->m_ran = false;
    invalidate_graph(h);
    hipStream_t s = h->stream;
    const int nf = h->nf;
    const int *fp = h->fac_in_f ? h->f_ptr.as<int>() : h->a_ptr.as<int>();
    const int *fi = h->fac_in_f ? h->f_ind.as<int>() : h->a_ind.as<int>();
    const double *fv = h->fac_in_f ? h->f_val.as<double>() : h->a_val.as<double>();
    hipEvent_t e0, e1;
    RC_HIP(hipEventCreate(&e0));
    RC_HIP(hipEventCreate(&e1));
    RC_HIP(hipEventRecord(e0, s));
    // rows of L = upper slices of the columns (dep_init = their lengths)
    RC_HIP(h->ic_up_ptr.ensure(((size_t)nf + 1) * sizeof(int)));
    RC_HIP(h->tmp_scan.ensure(scan_tmp_elems(nf + 3) * sizeof(int) + 64));
    scan_exclusive(h->dep_init.as<int>(), h->ic_up_ptr.as<int>(), nf, h->tmp_scan.as<int>(), s);
    int nup = 0;
    RC_HIP(hipMemcpyAsync(&nup, h->ic_up_ptr.as<int>() + nf, sizeof(int), hipMemcpyDeviceToHost, s));
    RC_HIP(hipStreamSynchronize(s));
    if ((int64_t)nup != h->nnz_low) return fail(h, RC_ERR_STATE, "IC(0): upper and lower slices differ in size");
    RC_HIP(h->ic_up_ind.ensure((size_t)std::max(nup, 1) * sizeof(int)));
    RC_HIP(h->ic_up_val.ensure((size_t)std::max(nup, 1) * sizeof(double)));
    RC_HIP(h->ic_l.ensure((size_t)std::max(nup, 1) * sizeof(double)));
    RC_HIP(h->ic_diag.ensure((size_t)nf * sizeof(double)));
    launch_ic0_upper(nf, fp, fi, fv, h->ic_up_ptr.as<int>(), h->ic_up_ind.as<int>(), h->ic_up_val.as<double>(),
                     h->ic_diag.as<double>(), s);
    // the elimination's output arrays
    if (h->pool_cap == 0) {
        h->pool_cap = 1024;
        h->scratch_cap = 8192;
    }
    h->out_cap = std::max<long long>(h->out_cap, h->nnz_low + 16);
    rc = alloc_factor_work(h);
    if (rc != RC_OK) return rc;
    RC_HIP(hipMemsetAsync(h->ctr.p, 0, kCtrCount * sizeof(unsigned long long), s));
    launch_ptr_to_start_cnt(nf, h->lo_ptr.as<int>(), h->out_start.as<long long>(), h->out_cnt.as<int>(), s);
    if (h->nnz_low)
        RC_HIP(hipMemcpyAsync(h->out_row.p, h->lo_row.p, (size_t)h->nnz_low * sizeof(int), hipMemcpyDeviceToDevice, s));
    launch_fill_sentinel<double>(h->diag64.as<double>(), nf, s);
    Ic0Args a;
    a.n = nf;
    a.n_compute = h->trim ? nf - 1 : nf;
    a.up_ptr = h->ic_up_ptr.as<int>();
    a.up_ind = h->ic_up_ind.as<int>();
    a.up_val = h->ic_up_val.as<double>();
    a.diag_a = h->ic_diag.as<double>();
    a.l = h->ic_l.as<double>();
    a.d = h->diag64.as<double>();
    a.level = h->level.as<int>();
    unsigned int *cw = reinterpret_cast<unsigned int *>(h->ctr.as<unsigned long long>() + kCtrCount - 3);
    a.claim = cw;
    a.status = cw + 2;
    a.spin_ticks = h->spin_ticks;
    launch_ic0(a, std::max(64, h->num_blocks * 4), s);
    launch_ic0_scatter(a.n_compute, a.up_ptr, a.up_ind, a.l, h->lo_ptr.as<int>(), h->lo_row.as<int>(),
                       h->out_val.as<double>(), s);
    RC_HIP(hipGetLastError());
    RC_HIP(hipEventRecord(e1, s));
    unsigned int status = 0;
    RC_HIP(hipMemcpyAsync(&status, a.status, sizeof(status), hipMemcpyDeviceToHost, s));
    unsigned long long top = (unsigned long long)h->nnz_low;
    RC_HIP(hipMemcpyAsync(h->ctr.as<unsigned long long>() + kCtrOutTop, &top, sizeof(top), hipMemcpyHostToDevice, s));
    RC_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (status & 1u) return fail(h, RC_ERR_TIMEOUT, "IC(0) row wait timed out");
    h->prec = h->opt.factor_precision;
    double ams = 0.0;
    rc = (h->prec == 32) ? assemble<float>(h, &ams) : assemble<double>(h, &ams);
    if (rc != RC_OK) return rc;
    h->fstats.n = nf;
    h->fstats.nnz_lower = h->nnz_low;
    h->fstats.nnz_factor = h->nnz_low + nf;
    h->fstats.nlevels = h->nlevels;
    h->fstats.factor_ms = ms;
    h->fstats.assemble_ms = ams;
    h->fstats.prepare_ms = prep_ms;
    // IC(0) reads the lower slice once and writes L once: 12 nnz_low + 4 (n+1)
    // in, (4+8) nnz_low + 8 n out
    h->fstats.bytes_factor = 24.0 * (double)h->nnz_low + 12.0 * (double)nf + 4.0;
    h->factored = true;
    if (st) *st = h->fstats;
    return RC_OK;
}

// ------------------------------------------------------------------ factor export
static int fetch_raw_factor(rc_handle_t h, std::vector<long long> &start, std::vector<int> &cnt, std::vector<int> &row,
                            std::vector<double> &val, std::vector<double> &diag)
{
    const size_t nf = (size_t)h->nf;
    start.resize(nf);
    cnt.resize(nf);
    diag.resize(nf);
    RC_HIP(hipMemcpyAsync(start.data(), h->out_start.p, nf * sizeof(long long), hipMemcpyDeviceToHost, h->stream));
    RC_HIP(hipMemcpyAsync(cnt.data(), h->out_cnt.p, nf * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    RC_HIP(hipMemcpyAsync(diag.data(), h->diag64.p, nf * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    unsigned long long top = 0;
    RC_HIP(hipMemcpyAsync(&top, h->ctr.as<unsigned long long>() + kCtrOutTop, sizeof(top), hipMemcpyDeviceToHost, h->stream));
    RC_HIP(hipStreamSynchronize(h->stream));
    row.resize((size_t)top);
    val.resize((size_t)top);
    if (top) {
        RC_HIP(hipMemcpyAsync(row.data(), h->out_row.p, top * sizeof(int), hipMemcpyDeviceToHost, h->stream));
        RC_HIP(hipMemcpyAsync(val.data(), h->out_val.p, top * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        RC_HIP(hipStreamSynchronize(h->stream));
    }
    return RC_OK;
}

static int rc_get_factor_impl(rc_handle_t h, int64_t *n, int64_t *nnz, int64_t *rowptr, int *colind, double *val, double *diag)
{
    RC_CHECK_HANDLE();
    if (!h->factored) return fail(h, RC_ERR_STATE, "no factor");
    const int ns = h->ns;
    if (n) *n = ns;
    if (nnz) *nnz = h->nnz_lt + ns;
    if (!rowptr && !colind && !val && !diag) return RC_OK;
    std::vector<int> ptr((size_t)ns + 1), ind((size_t)h->nnz_lt);
    std::vector<double> v((size_t)h->nnz_lt), d((size_t)ns);
    RC_HIP(hipMemcpyAsync(ptr.data(), h->lt_ptr.p, ptr.size() * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    if (h->nnz_lt)
        RC_HIP(hipMemcpyAsync(ind.data(), h->lt_ind.p, ind.size() * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    if (h->prec == 64) {
        if (h->nnz_lt) RC_HIP(hipMemcpyAsync(v.data(), h->lt_val.p, v.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        RC_HIP(hipMemcpyAsync(d.data(), h->diagT.p, d.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        RC_HIP(hipStreamSynchronize(h->stream));
    } else {
        std::vector<float> vf(v.size()), df(d.size());
        if (h->nnz_lt) RC_HIP(hipMemcpyAsync(vf.data(), h->lt_val.p, vf.size() * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        RC_HIP(hipMemcpyAsync(df.data(), h->diagT.p, df.size() * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        RC_HIP(hipStreamSynchronize(h->stream));
        for (size_t i = 0; i < v.size(); ++i) v[i] = vf[i];
        for (size_t i = 0; i < d.size(); ++i) d[i] = df[i];
    }
    int64_t q = 0;
    if (rowptr) rowptr[0] = 0;
    for (int k = 0; k < ns; ++k) {
        if (colind) colind[q] = k;                     // unit diagonal first
        if (val) val[q] = 1.0;
        ++q;
        for (int e = ptr[(size_t)k]; e < ptr[(size_t)k + 1]; ++e, ++q) {
            if (colind) colind[q] = ind[(size_t)e];
            if (val) val[q] = v[(size_t)e];
        }
        if (rowptr) rowptr[k + 1] = q;
    }
    if (diag) std::memcpy(diag, d.data(), d.size() * sizeof(double));
    return RC_OK;
}

static int rc_dump_factor_impl(rc_handle_t h, const char *prefix)
{
    RC_CHECK_HANDLE();
    if (!prefix) return fail(h, RC_ERR_ARG, "null prefix");
    if (!h->factored) return fail(h, RC_ERR_STATE, "no factor");
    std::vector<long long> start;
    std::vector<int> cnt, row;
    std::vector<double> val, diag;
    int rc = fetch_raw_factor(h, start, cnt, row, val, diag);
    if (rc != RC_OK) return rc;
    // CSC of L with the unit diagonal first (driver_local.cpp:881-934, 969-983)
    const int nf = h->nf;
    std::vector<int64_t> cp((size_t)nf + 1, 0);
    for (int k = 0; k < nf; ++k) cp[(size_t)k + 1] = cp[(size_t)k] + cnt[(size_t)k] + 1;
    std::vector<int> ri((size_t)cp[(size_t)nf]);
    std::vector<double> vv((size_t)cp[(size_t)nf]);
    for (int k = 0; k < nf; ++k) {
        int64_t q = cp[(size_t)k];
        ri[(size_t)q] = k;
        vv[(size_t)q] = 1.0;
        ++q;
        for (int t = 0; t < cnt[(size_t)k]; ++t, ++q) {
            ri[(size_t)q] = row[(size_t)(start[(size_t)k] + t)];
            vv[(size_t)q] = val[(size_t)(start[(size_t)k] + t)];
        }
    }
    std::string p(prefix);
    rc = write_mtx_csc(p + "l_sol.mtx", nf, nf, cp.data(), ri.data(), vv.data(), h->err);
    if (rc != RC_OK) return rc;
    FILE *f = fopen((p + "sol_diag.txt").c_str(), "wb");
    if (!f) return fail(h, RC_ERR_IO, "cannot write " + p + "sol_diag.txt");
    for (int k = 0; k < nf; ++k) fprintf(f, "%.17g\n", diag[(size_t)k]);
    fclose(f);
    return RC_OK;
}

static int rc_load_factor_impl(rc_handle_t h, const char *prefix)
{
    RC_CHECK_HANDLE();
    if (h->host_only) return fail(h, RC_ERR_NO_DEVICE, "host-only handle");
    if (!prefix) return fail(h, RC_ERR_ARG, "null prefix");
    RC_HIP(hipSetDevice(h->opt.device));
    int rc = prepare(h);
    if (rc != RC_OK) return rc;
    std::string p(prefix);
    MtxTriplets t;
    rc = read_mtx(p + "l_sol.mtx", t, h->err, h->opt.host_threads);
    if (rc != RC_OK) return rc;
    const int nf = h->nf;
    if (t.nrows != nf || t.ncols != nf) return fail(h, RC_ERR_FORMAT, "factor dimension does not match the system");
    std::vector<double> diag;
    {
        std::ifstream in(p + "sol_diag.txt");
        if (!in) return fail(h, RC_ERR_IO, "cannot open " + p + "sol_diag.txt");
        double v;
        while (in >> v) diag.push_back(v);
    }
    if ((int)diag.size() != nf) return fail(h, RC_ERR_FORMAT, "sol_diag.txt length does not match");
    // columns of L, strict lower entries, sorted by row
    std::vector<int> cnt((size_t)nf, 0);
    for (size_t e = 0; e < t.rows.size(); ++e) {
        if (t.rows[e] < t.cols[e]) return fail(h, RC_ERR_FORMAT, "factor entry above the diagonal");
        if (t.rows[e] > t.cols[e]) cnt[(size_t)t.cols[e]]++;
    }
    std::vector<long long> start((size_t)nf, 0);
    long long tot = 0;
    for (int k = 0; k < nf; ++k) { start[(size_t)k] = tot; tot += cnt[(size_t)k]; }
    if (tot > kMaxFactorEntries) return fail(h, RC_ERR_FORMAT, "factor has more than 2^31 - 2 off-diagonal entries");
    std::vector<int> row((size_t)tot);
    std::vector<double> val((size_t)tot);
    std::vector<long long> pos(start);
    for (size_t e = 0; e < t.rows.size(); ++e)
        if (t.rows[e] > t.cols[e]) {
            long long q = pos[(size_t)t.cols[e]]++;
            row[(size_t)q] = t.rows[e];
            val[(size_t)q] = t.vals[e];
        }
    std::vector<int> level((size_t)nf, 0);
    for (int k = 0; k < nf; ++k) {
        std::vector<std::pair<int, double>> col((size_t)cnt[(size_t)k]);
        for (int i = 0; i < cnt[(size_t)k]; ++i) col[(size_t)i] = {row[(size_t)(start[(size_t)k] + i)], val[(size_t)(start[(size_t)k] + i)]};
        std::sort(col.begin(), col.end());
        for (int i = 0; i < cnt[(size_t)k]; ++i) {
            row[(size_t)(start[(size_t)k] + i)] = col[(size_t)i].first;
            val[(size_t)(start[(size_t)k] + i)] = col[(size_t)i].second;
            int r = col[(size_t)i].first;
            level[(size_t)r] = std::max(level[(size_t)r], level[(size_t)k] + 1);
        }
    }
    if (h->pool_cap == 0) {
        h->pool_cap = std::max<long long>(1024, (long long)std::ceil(h->opt.pool_factor * (double)h->nnz_low));
        h->scratch_cap = 8192;
    }
    h->out_cap = std::max<long long>(h->out_cap, tot + 16);
    rc = alloc_factor_work(h);
    if (rc != RC_OK) return rc;
    RC_HIP(hipMemcpyAsync(h->out_start.p, start.data(), (size_t)nf * sizeof(long long), hipMemcpyHostToDevice, h->stream));
    RC_HIP(hipMemcpyAsync(h->out_cnt.p, cnt.data(), (size_t)nf * sizeof(int), hipMemcpyHostToDevice, h->stream));
    RC_HIP(hipMemcpyAsync(h->diag64.p, diag.data(), (size_t)nf * sizeof(double), hipMemcpyHostToDevice, h->stream));
    RC_HIP(hipMemcpyAsync(h->level.p, level.data(), (size_t)nf * sizeof(int), hipMemcpyHostToDevice, h->stream));
    if (tot) {
        RC_HIP(hipMemcpyAsync(h->out_row.p, row.data(), (size_t)tot * sizeof(int), hipMemcpyHostToDevice, h->stream));
        RC_HIP(hipMemcpyAsync(h->out_val.p, val.data(), (size_t)tot * sizeof(double), hipMemcpyHostToDevice, h->stream));
    }
    unsigned long long ctr[kCtrCount] = {0};
    ctr[kCtrOutTop] = (unsigned long long)tot;
    RC_HIP(hipMemcpyAsync(h->ctr.p, ctr, sizeof(ctr), hipMemcpyHostToDevice, h->stream));
    RC_HIP(hipStreamSynchronize(h->stream));
    std::memset(&h->fstats, 0, sizeof(h->fstats));
    h->fstats.nnz_factor = tot + nf;
    int64_t fills = 0, maxd = 0;
    int zp = 0;
    for (int k = 0; k < nf; ++k) {
        fills += std::max(cnt[(size_t)k] - 1, 0);
        maxd = std::max<int64_t>(maxd, cnt[(size_t)k]);
        if (diag[(size_t)k] == 0.0) ++zp;
    }
    h->fstats.fills = fills;
    h->fstats.max_degree = maxd;
    h->fstats.zero_pivots = zp;
    h->prec = h->opt.factor_precision;
    double ams = 0.0;
    invalidate_graph(h);
    rc = (h->prec == 32) ? assemble<float>(h, &ams) : assemble<double>(h, &ams);
    if (rc != RC_OK) return rc;
    h->fstats.n = nf;
    h->fstats.nnz_lower = h->nnz_low;
    h->fstats.nlevels = h->nlevels;
    h->fstats.assemble_ms = ams;
    h->factored = true;
    return RC_OK;
}

// ------------------------------------------------------------------ solve
}  // extern "C"

SolveArgs solve_args(rc_handle_t h)
{
    SolveArgs a;
    a.n = h->ns;
    a.ntiles = h->ntiles;
    a.a_tile_ord = h->tile_ord.as<int>();
    a.a_off = h->a_off.as<int>();
    a.a_width = h->a_width.as<int>();
    a.a_ind = h->a_ell_ind.as<int>();
    a.a_val = h->a_ell_val.as<double>();
    a.nchunks = h->nchunks;
    a.chunk_start = h->chunk_start.as<int>();
    a.chunk_cnt = h->chunk_cnt.as<int>();
    a.cm_f = h->cm_f.as<int4>();
    a.cm_b = h->cm_b.as<int4>();
    a.lf_off = h->lf_off.as<int>();
    a.lf_width = h->lf_width.as<int>();
    a.lf_ind = h->lf_ind.as<int>();
    a.lf_zero = (int)h->ell_l;
    a.lb_zero = (int)h->ell_u;
    a.lb_off = h->lb_off.as<int>();
    a.lb_width = h->lb_width.as<int>();
    a.lb_ind = h->lb_ind.as<int>();
    a.lf_val = h->lf_val.p;
    a.lb_val = h->lb_val.p;
    a.diag = h->diagP.p;
    a.x = h->x.as<double>();
    a.r = h->r.as<double>();
    a.ap = h->ap.as<double>();
    a.p0 = h->p0.as<double>();
    a.p1 = h->p1.as<double>();
    a.y = h->y.p;
    a.z = h->z.p;
    a.sc = h->scal.as<PcgScalars>();
    a.partials = h->partials.as<double>();
    a.ctr_fwd = h->ctrs.as<unsigned int>();
    a.ctr_bwd = h->ctrs.as<unsigned int>() + kClaimQueues * kClaimStride;
    a.reduce_blocks = h->reduce_blocks;
    a.sweep_blocks = h->sweep_blocks;
    a.bulk_chunks = h->bulk_chunks;
    a.tail_level = h->tail_level;
    a.tail_pos0 = h->tail_pos0;
    a.ntail = h->ntail;
    a.ntail_levels = h->ntail_levels;
    a.tail_lds_bytes = h->tail_lds_bytes;
    a.tl_start = h->tl_start.as<int>();
    a.tl_cnt = h->tl_cnt.as<int>();
    a.tf_off = h->tf_off.as<int>();
    a.tf_slices = h->tf_slices.as<int>();
    a.tf_ind = h->tf_ind.as<int>();
    a.tf_val = h->tf_val.p;
    a.tb_off = h->tb_off.as<int>();
    a.tb_slices = h->tb_slices.as<int>();
    a.tb_ind = h->tb_ind.as<int>();
    a.tb_val = h->tb_val.p;
    a.tp_ptr = h->tp_ptr.as<int>();
    a.tp_ind = h->tp_ind.as<int>();
    a.tp_val = h->tp_val.p;
    a.tacc = h->tacc.p;
    a.nlevels = h->nlevels;
    a.spin_ticks = h->spin_ticks;
    a.b = h->b.as<double>();
    a.zero = 0;
    a.nap_max = h->nap_max;
    a.sys_of = nullptr;
    a.alpha_sys = nullptr;
    if (h->nsys > 0) {
        BatchArgs b;
        b.nsys = h->nsys;
        b.nblk = h->nblk;
        batch_carve(b, h->bt_state.p);
        a.sys_of = h->sys_pos.as<int>();
        a.alpha_sys = b.alpha;
    }
    return a;
}

extern "C" {

static BatchArgs batch_args(rc_handle_t h)
{
    BatchArgs b;
    b.nsys = h->nsys;
    b.nblk = h->nblk;
    batch_carve(b, h->bt_state.p);
    b.sys_of = h->sys_pos.as<int>();
    b.spos = h->spos.as<int>();
    const int *tab = h->blk_tab.as<int>();
    b.blk_lo = tab;
    b.blk_hi = tab + h->nblk;
    b.blk_sys = tab + 2 * h->nblk;
    b.sys_blk0 = tab + 3 * h->nblk;
    return b;
}

static int alloc_solve(rc_handle_t h)
{
    const size_t n = (size_t)h->ns + 1;          // + dummy slot
    const size_t s = h->prec == 32 ? 4 : 8;
    const size_t old_y = h->y.p ? (size_t)h->y.p : 0;
    RC_HIP(h->x.ensure(n * 8));
    RC_HIP(h->r.ensure(n * 8));
    RC_HIP(h->ap.ensure(n * 8));
    RC_HIP(h->p0.ensure(n * 8));
    RC_HIP(h->p1.ensure(n * 8));
    RC_HIP(h->b.ensure(n * 8));
    RC_HIP(h->bu.ensure(n * 8));
    RC_HIP(h->y.ensure(n * s));
    RC_HIP(h->z.ensure(n * s));
    RC_HIP(h->scal.ensure(sizeof(PcgScalars)));
    RC_HIP(h->partials.ensure((size_t)h->reduce_blocks * 2 * sizeof(double)));
    RC_HIP(h->ctrs.ensure(kClaimWords * sizeof(unsigned int)));
    if ((size_t)h->y.p != old_y) invalidate_graph(h);
    // the dummy slot n of every vector holds 0 from here on: ELL padding and
    // the zero entries that masked lanes load point at it, also when the
    // sweeps run before any PCG (timed or tuned first) on memory whose last
    // user left a sentinel there
    const size_t ns = (size_t)h->ns;
    for (DevVec *v : {&h->x, &h->r, &h->ap, &h->p0, &h->p1, &h->b})
        RC_HIP(hipMemsetAsync((char *)v->p + ns * 8, 0, 8, h->stream));
    RC_HIP(hipMemsetAsync((char *)h->y.p + ns * s, 0, s, h->stream));
    RC_HIP(hipMemsetAsync((char *)h->z.p + ns * s, 0, s, h->stream));
    return RC_OK;
}

static int rc_make_rhs_impl(rc_handle_t h, int kind, uint64_t seed, double *b)
{
    RC_CHECK_HANDLE();
    if (h->kind < 0 || !b) return fail(h, RC_ERR_ARG, "no matrix or null array");
    const int n = h->n_user;
    if (kind == RC_RHS_ZERO_SUM_UNIFORM) {                  // solver.hpp:43-61
        std::mt19937 gen((unsigned long)seed);
        std::uniform_real_distribution<double> dist(0.0, 1.0);
        double sum = 0.0;
        for (int i = 0; i < n; ++i) { b[i] = dist(gen); sum += b[i]; }
        double mean = sum / n;
        for (int i = 0; i < n; ++i) b[i] -= mean;
        return RC_OK;
    }
    if (kind != RC_RHS_NORMAL && kind != RC_RHS_NORMAL_PROJECTED) return fail(h, RC_ERR_ARG, "unknown rhs kind");
    for (int i = 0; i < n; ++i) {
        U4 o = philox_keyed(seed, (uint32_t)i, 0u, kTagNormal);
        double u1 = u01_open0(o.x, o.y), u2 = u01_open1(o.z, o.w);
        b[i] = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    if (kind == RC_RHS_NORMAL_PROJECTED) {
        Csc user;
        if (h->kind == RC_INPUT_GROUNDED) trim_last(h->base, user);
        const Csc &g = h->kind == RC_INPUT_GROUNDED ? user : h->base;
        std::vector<int> comp;
        int nc = components(g, comp);
        std::vector<double> s((size_t)nc, 0.0), c((size_t)nc, 0.0);
        for (int i = 0; i < n; ++i) { s[(size_t)comp[(size_t)i]] += b[i]; c[(size_t)comp[(size_t)i]] += 1.0; }
        for (int i = 0; i < n; ++i) b[i] -= s[(size_t)comp[(size_t)i]] / c[(size_t)comp[(size_t)i]];
    }
    return RC_OK;
}

static int rc_set_reference_solution_impl(rc_handle_t h, const double *x_ref)
{
    RC_CHECK_HANDLE();
    if (!x_ref) { h->xref.clear(); return RC_OK; }
    if (h->kind < 0) return fail(h, RC_ERR_STATE, "no matrix loaded");
    h->xref.assign(x_ref, x_ref + h->n_user);
    return RC_OK;
}

static int set_rhs_device(rc_handle_t h, const double *b_user)
{
    int rc = alloc_solve(h);
    if (rc != RC_OK) return rc;
    RC_HIP(hipMemcpyAsync(h->bu.p, b_user, (size_t)h->ns * sizeof(double), hipMemcpyHostToDevice, h->stream));
    launch_gather(h->bu.as<double>(), h->idx_user.as<int>(), h->b.as<double>(), h->ns, h->stream);
    RC_HIP(hipGetLastError());
    h->rhs_set = true;
    h->rhs_mode = 1;
    return RC_OK;
}

// the single-system / batch PCG run (defined with the other PCG runs below)
static int solve_resident(rc_handle_t h, double tol, rc_solve_stats *st);

static int fetch_solution(rc_handle_t h, double *x_user)
{
    // x_user[idx_user[p]] = x[p]
    launch_scatter(h->x.as<double>(), h->idx_user.as<int>(), h->bu.as<double>(), h->ns, h->stream);
    RC_HIP(hipMemcpyAsync(x_user, h->bu.p, (size_t)h->ns * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    RC_HIP(hipStreamSynchronize(h->stream));
    return RC_OK;
}

static double a_norm_err(rc_handle_t h, const double *x)
{
    // ||x - xref||_A / ||xref||_A on the user system (host, diagnostics only)
    Csc user;
    if (h->kind == RC_INPUT_GROUNDED) trim_last(h->base, user);
    const Csc &g = h->kind == RC_INPUT_GROUNDED ? user : h->base;
    const int n = h->n_user;
    std::vector<double> d((size_t)n);
    for (int i = 0; i < n; ++i) d[(size_t)i] = x[i] - h->xref[(size_t)i];
    double num = 0.0, den = 0.0;
    for (int j = 0; j < n; ++j) {
        double ad = 0.0, ax = 0.0;
        for (int64_t p = g.ptr[(size_t)j]; p < g.ptr[(size_t)j + 1]; ++p) {
            ad += g.val[(size_t)p] * d[(size_t)g.ind[(size_t)p]];
            ax += g.val[(size_t)p] * h->xref[(size_t)g.ind[(size_t)p]];
        }
        num += d[(size_t)j] * ad;
        den += h->xref[(size_t)j] * ax;
    }
    return den > 0.0 ? std::sqrt(std::max(num, 0.0) / den) : -1.0;
}

static int rc_solve_impl(rc_handle_t h, const double *b, double *x, double tol, rc_solve_stats *st)
{
    RC_CHECK_HANDLE();
    if (h->host_only) return fail(h, RC_ERR_NO_DEVICE, "host-only handle");
    if (!b || !x || !(tol >= 0.0)) return fail(h, RC_ERR_ARG, "bad solve arguments");
    if (!h->factored) return fail(h, RC_ERR_STATE, "rc_factor must precede rc_solve");
    RC_HIP(hipSetDevice(h->opt.device));
    int rc = set_rhs_device(h, b);
    if (rc != RC_OK) return rc;
    rc_solve_stats local;
    rc_solve_stats *sp = st ? st : &local;
    rc = solve_resident(h, tol, sp);
    if (rc != RC_OK && rc != RC_ERR_BREAKDOWN) return rc;
    int rc2 = fetch_solution(h, x);
    if (rc2 != RC_OK) return rc2;
    if (!h->xref.empty()) sp->err_vs_oracle = a_norm_err(h, x);
    return rc;
}

static int rc_solve_batch_impl(rc_handle_t h, const double *b, double *x, double tol, rc_solve_stats *per_system)
{
    RC_CHECK_HANDLE();
    if (h->nsys <= 0) return fail(h, RC_ERR_STATE, "handle holds no batch (rc_set_batch)");
    rc_solve_stats total;
    int rc = rc_solve(h, b, x, tol, &total);
    if (rc != RC_OK && rc != RC_ERR_BREAKDOWN) return rc;
    if (per_system && h->bstats.size() == (size_t)h->nsys)
        std::memcpy(per_system, h->bstats.data(), h->bstats.size() * sizeof(rc_solve_stats));
    return rc;
}

static int rc_get_batch_stats_impl(rc_handle_t h, rc_solve_stats *per_system)
{
    RC_CHECK_HANDLE();
    if (h->nsys <= 0 || h->bstats.size() != (size_t)h->nsys) return fail(h, RC_ERR_STATE, "no batched solve");
    if (!per_system) return fail(h, RC_ERR_ARG, "null stats array");
    std::memcpy(per_system, h->bstats.data(), h->bstats.size() * sizeof(rc_solve_stats));
    return RC_OK;
}

// ------------------------------------------------------------------ multi-RHS
}  // extern "C"

template <typename T>
static int build_multi(rc_handle_t h)
{
    if (h->multi_ready) return RC_OK;
    hipStream_t s = h->stream;
    const int ns = h->ns;
    const size_t nsz = (size_t)ns + 1;
    auto pos_csr = [&](const int *src_ptr, EllSrc src, DevVec &ptr, DevVec &ind, DevVec &val, size_t vbytes) -> int {
        RC_HIP(ptr.ensure(nsz * sizeof(int)));
        launch_pos_csr_count(ns, h->order.as<int>(), src_ptr, h->tmp_a.as<int>(), s);
        scan_exclusive(h->tmp_a.as<int>(), ptr.as<int>(), ns, h->tmp_scan.as<int>(), s);
        int nnz = 0;
        RC_HIP(hipMemcpyAsync(&nnz, ptr.as<int>() + ns, sizeof(int), hipMemcpyDeviceToHost, s));
        RC_HIP(hipStreamSynchronize(s));
        RC_HIP(ind.ensure((size_t)std::max(nnz, 1) * sizeof(int)));
        RC_HIP(val.ensure((size_t)std::max(nnz, 1) * vbytes));
        return RC_OK;
    };
    EllSrc as{h->a_ptr.as<int>(), h->a_ind.as<int>(), h->a_val.p, nullptr, nullptr, h->pos_of.as<int>()};
    int rc = pos_csr(h->a_ptr.as<int>(), as, h->m_aptr, h->m_aind, h->m_aval, sizeof(double));
    if (rc != RC_OK) return rc;
    launch_pos_csr_fill<double, double>(ns, h->order.as<int>(), as, h->m_aptr.as<int>(), h->m_aind.as<int>(),
                                        h->m_aval.as<double>(), s);
    EllSrc fs{h->l_ptr.as<int>(), nullptr, h->lt_val.p, h->vals0.as<int>(), h->rowk.as<int>(), nullptr};
    rc = pos_csr(h->l_ptr.as<int>(), fs, h->m_lptr, h->m_lind, h->m_lval, sizeof(T));
    if (rc != RC_OK) return rc;
    launch_pos_csr_fill<T, T>(ns, h->order.as<int>(), fs, h->m_lptr.as<int>(), h->m_lind.as<int>(),
                              h->m_lval.as<T>(), s);
    EllSrc bs{h->lt_ptr.as<int>(), h->lt_ind.as<int>(), h->lt_val.p, nullptr, nullptr, h->pos_of.as<int>()};
    rc = pos_csr(h->lt_ptr.as<int>(), bs, h->m_ltptr, h->m_ltind, h->m_ltval, sizeof(T));
    if (rc != RC_OK) return rc;
    launch_pos_csr_fill<T, T>(ns, h->order.as<int>(), bs, h->m_ltptr.as<int>(), h->m_ltind.as<int>(),
                              h->m_ltval.as<T>(), s);
    RC_HIP(hipGetLastError());
    RC_HIP(hipStreamSynchronize(s));
    h->multi_ready = true;
    return RC_OK;
}

extern "C" {

static int multi_args(rc_handle_t h, int k, MultiArgs &a)
{
    int K = 1;
    while (K < k) K <<= 1;
    const size_t n = (size_t)h->ns;
    const size_t tot = (n + 1) * (size_t)K;
    const size_t sz = h->prec == 32 ? 4 : 8;
    RC_HIP(h->m_x.ensure(tot * 8));
    RC_HIP(h->m_r.ensure(tot * 8));
    RC_HIP(h->m_ap.ensure(tot * 8));
    RC_HIP(h->m_p0.ensure(tot * 8));
    RC_HIP(h->m_p1.ensure(tot * 8));
    RC_HIP(h->m_b.ensure(tot * 8));
    RC_HIP(h->m_y.ensure(tot * sz));
    RC_HIP(h->m_z.ensure(tot * sz));
    RC_HIP(h->m_sc.ensure(sizeof(MultiScalars)));
    RC_HIP(h->m_part.ensure((size_t)h->reduce_blocks * 2 * K * sizeof(double)));
    RC_HIP(h->ctrs.ensure(kClaimWords * sizeof(unsigned int)));
    RC_HIP(h->pos_of_user.ensure((n + 1) * sizeof(int)));
    launch_inverse_perm((int)n, h->idx_user.as<int>(), h->pos_of_user.as<int>(), h->stream);
    a.n = (int)n;
    a.K = K;
    a.k = k;
    a.a_ptr = h->m_aptr.as<int>();
    a.a_ind = h->m_aind.as<int>();
    a.a_val = h->m_aval.as<double>();
    a.row_ord = h->pos_of_user.as<int>();
    a.l_ptr = h->m_lptr.as<int>();
    a.l_ind = h->m_lind.as<int>();
    a.lt_ptr = h->m_ltptr.as<int>();
    a.lt_ind = h->m_ltind.as<int>();
    a.l_val = h->m_lval.p;
    a.lt_val = h->m_ltval.p;
    a.diag = h->diagP.p;
    a.nchunks = h->nchunks;
    a.chunk_start = h->chunk_start.as<int>();
    a.chunk_cnt = h->chunk_cnt.as<int>();
    a.x = h->m_x.as<double>();
    a.r = h->m_r.as<double>();
    a.ap = h->m_ap.as<double>();
    a.p0 = h->m_p0.as<double>();
    a.p1 = h->m_p1.as<double>();
    a.y = h->m_y.p;
    a.z = h->m_z.p;
    a.b = h->m_b.as<double>();
    a.sc = h->m_sc.as<MultiScalars>();
    a.partials = h->m_part.as<double>();
    a.ctr_fwd = h->ctrs.as<unsigned int>();
    a.ctr_bwd = h->ctrs.as<unsigned int>() + kClaimQueues * kClaimStride;
    a.reduce_blocks = h->reduce_blocks;
    a.sweep_blocks = h->sweep_blocks;
    a.spin_ticks = h->spin_ticks;
    return RC_OK;
}

static int rc_set_rhs_multi_impl(rc_handle_t h, int k, const double *B)
{
    RC_CHECK_HANDLE();
    if (h->host_only) return fail(h, RC_ERR_NO_DEVICE, "host-only handle");
    if (k <= 0 || k > kMaxRhs || !B) return fail(h, RC_ERR_ARG, "bad multi-RHS arguments");
    if (h->kind < 0) return fail(h, RC_ERR_STATE, "no matrix loaded");
    if (h->nsys > 0) return fail(h, RC_ERR_STATE, "a batch is solved with rc_solve_batch (one RHS per system)");
    RC_HIP(hipSetDevice(h->opt.device));
    const size_t n = (size_t)h->n_user;
    RC_HIP(h->m_B.ensure(n * (size_t)k * 8));
    RC_HIP(hipMemcpyAsync(h->m_B.p, B, n * (size_t)k * 8, hipMemcpyHostToDevice, h->stream));
    RC_HIP(hipStreamSynchronize(h->stream));
    h->m_k = k;
    h->m_n = (int)n;
    h->m_solved = false;
    h->m_ran = false;
    h->rhs_mode = 2;
    return RC_OK;
}

// ------------------------------------------------------------------ PCG runs
// One run is the PCG of one handle in one of two modes: the single-RHS solve
// (one system, or one right-hand side per system of a batch) or the
// multi-RHS solve.  A run is split into
//   begin: scalars uploaded and the init kernels launched on the handle's stream;
//   pairs: two iterations at a time, captured into a hipGraph (cached per
//          handle and mode for a run of one handle; one graph holding every
//          member's pair for a group of handles);
//   end:   the true residual and the statistics.
// Kernels of a finished run (done flag set) return at once, so the pairs of
// a group keep launching until every member is done.
enum { kRunSingle = 1, kRunMulti = 2 };

static int run_state_ok(rc_handle_t h, int mode)
{
    if (h->host_only) return fail(h, RC_ERR_NO_DEVICE, "host-only handle");
    if (mode == kRunSingle) {
        if (!h->factored || !h->rhs_set) return fail(h, RC_ERR_STATE, "factor and rc_set_rhs must precede");
        return RC_OK;
    }
    if (!h->factored || h->m_k <= 0) return fail(h, RC_ERR_STATE, "rc_factor and rc_set_rhs_multi must precede");
    if (h->nsys > 0) return fail(h, RC_ERR_STATE, "a batch is solved with rc_solve_batch (one RHS per system)");
    if ((size_t)h->ns != (size_t)h->n_user || h->m_n != h->n_user)
        return fail(h, RC_ERR_STATE, "right-hand sides do not match the system");
    return RC_OK;
}

static int run_begin(rc_handle_t h, int mode, double tol, MultiArgs &ma)
{
    hipStream_t s = h->stream;
    if (mode == kRunSingle) {
        int rc = alloc_solve(h);
        if (rc != RC_OK) return rc;
        PcgScalars sc;
        std::memset(&sc, 0, sizeof(sc));
        sc.tol = tol;
        sc.maxit = h->opt.maxit;
        sc.stop_rule = h->opt.stop_rule;
        SolveArgs a = solve_args(h);
        RC_HIP(hipMemcpyAsync(h->scal.p, &sc, sizeof(sc), hipMemcpyHostToDevice, s));
        if (h->nsys > 0) {
            BatchArgs b = batch_args(h);
            RC_HIP(hipMemsetAsync(h->bt_state.p, 0, batch_state_bytes(h->nsys, h->nblk), s));
            if (h->prec == 32) launch_batch_init<float>(a, b, s);
            else launch_batch_init<double>(a, b, s);
        } else if (h->prec == 32) {
            launch_pcg_init<float>(a, s);
        } else {
            launch_pcg_init<double>(a, s);
        }
        RC_HIP(hipGetLastError());
        return RC_OK;
    }
    int rc = h->prec == 32 ? build_multi<float>(h) : build_multi<double>(h);
    if (rc != RC_OK) return rc;
    rc = multi_args(h, h->m_k, ma);
    if (rc != RC_OK) return rc;
    MultiScalars sc;
    std::memset(&sc, 0, sizeof(sc));
    sc.tol = tol;
    sc.maxit = h->opt.maxit;
    sc.stop_rule = h->opt.stop_rule;
    RC_HIP(hipMemcpyAsync(h->m_sc.p, &sc, sizeof(sc), hipMemcpyHostToDevice, s));
    launch_multi_gather(h->m_B.as<double>(), h->idx_user.as<int>(), ma.n, ma.k, ma.K, h->m_b.as<double>(), s);
    if (h->prec == 32) launch_multi_init<float>(ma, s);
    else launch_multi_init<double>(ma, s);
    RC_HIP(hipGetLastError());
    return RC_OK;
}

// Two iterations of the run on stream s (parities 0 and 1).
static void run_pair(rc_handle_t h, int mode, const MultiArgs &ma, hipStream_t s)
{
    if (mode == kRunMulti) {
        for (int par = 0; par < 2; ++par) {
            if (h->prec == 32) launch_multi_iteration<float>(ma, par, s);
            else launch_multi_iteration<double>(ma, par, s);
        }
        return;
    }
    SolveArgs a = solve_args(h);
    for (int par = 0; par < 2; ++par) {
        if (h->nsys > 0) {
            BatchArgs b = batch_args(h);
            if (h->prec == 32) launch_batch_iteration<float>(a, b, par, s);
            else launch_batch_iteration<double>(a, b, par, s);
        } else if (h->prec == 32) {
            launch_pcg_iteration<float>(a, par, s);
        } else {
            launch_pcg_iteration<double>(a, par, s);
        }
    }
}

static const int *run_done_flag(rc_handle_t h, int mode)
{
    if (mode == kRunMulti) return reinterpret_cast<const int *>(h->m_sc.as<char>() + offsetof(MultiScalars, all_done));
    return reinterpret_cast<const int *>(h->scal.as<char>() + offsetof(PcgScalars, done));
}

// The cached graph of a run of one handle: two iterations, rebuilt when the
// buffers it was captured with change.
static int run_graph(rc_handle_t h, int mode, const MultiArgs &ma, hipGraphExec_t *out)
{
    const void *key = mode == kRunMulti ? h->m_x.p : h->y.p;
    hipGraphExec_t *slot = mode == kRunMulti ? &h->m_graph : &h->graph;
    const int kk = mode == kRunMulti ? ma.K : 0;
    if (*slot && h->graph_key[mode] == key && h->graph_prec_of[mode] == h->prec && h->graph_K[mode] == kk) {
        *out = *slot;
        return RC_OK;
    }
    if (*slot) (void)hipGraphExecDestroy(*slot);
    *slot = nullptr;
    hipGraph_t g;
    RC_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    run_pair(h, mode, ma, h->stream);
    RC_HIP(hipStreamEndCapture(h->stream, &g));
    hipError_t e = hipGraphInstantiate(slot, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        *slot = nullptr;
        return fail(h, RC_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    h->graph_key[mode] = key;
    h->graph_prec_of[mode] = h->prec;
    h->graph_K[mode] = kk;
    *out = *slot;
    return RC_OK;
}

// Residual and statistics of a finished run; st holds 1 entry (single mode:
// the batch's worst system for a batch) or k entries (multi mode).  Sets
// *broke when a breakdown ended the run (some RHS / system of it).
static int run_end(rc_handle_t h, int mode, double tol, const MultiArgs &ma, float ms_setup, float ms_pcg,
                   rc_solve_stats *st, bool *broke)
{
    hipStream_t s = h->stream;
    *broke = false;
    if (mode == kRunMulti) {
        launch_multi_residual(ma, s);
        MultiScalars sc;
        RC_HIP(hipMemcpyAsync(&sc, h->m_sc.p, sizeof(sc), hipMemcpyDeviceToHost, s));
        RC_HIP(hipStreamSynchronize(s));
        h->m_solved = true;
        h->m_ran = true;
        const int k = h->m_k;
        h->m_stats.assign((size_t)k, rc_solve_stats{});
        for (int q = 0; q < k; ++q) {
            rc_solve_stats &o = h->m_stats[(size_t)q];
            *broke = *broke || sc.breakdown[q];
            o.iterations = sc.iter[q];
            const bool met = h->opt.stop_rule == RC_STOP_REFERENCE
                                 ? (sc.rz0[q] <= tol || std::sqrt(sc.rz[q]) / std::sqrt(sc.rz0[q]) <= tol)
                                 : (sc.bb[q] <= 0.0 || std::sqrt(sc.rr[q]) / std::sqrt(sc.bb[q]) <= tol);
            o.converged = (sc.done[q] && !sc.breakdown[q] && (sc.iter[q] < h->opt.maxit || met)) ? 1 : 0;
            o.rel_res_true = sc.bb[q] > 0.0 ? std::sqrt(sc.res2[q]) / std::sqrt(sc.bb[q]) : 0.0;
            o.rel_res_precond = sc.rz0[q] > 0.0 ? std::sqrt(std::max(sc.rz[q], 0.0)) / std::sqrt(sc.rz0[q]) : 0.0;
            o.pcg_ms = ms_pcg;
            o.setup_ms = ms_setup;
            o.err_vs_oracle = -1.0;
        }
        if (st) std::memcpy(st, h->m_stats.data(), (size_t)k * sizeof(rc_solve_stats));
        if (sc.status & kPcgTimeout) return fail(h, RC_ERR_TIMEOUT, "triangular solve wait timed out");
        if (*broke) return fail(h, RC_ERR_BREAKDOWN, "PCG breakdown in at least one right-hand side");
        return RC_OK;
    }
    SolveArgs a = solve_args(h);
    if (h->nsys > 0) launch_batch_residual(a, batch_args(h), s);
    else launch_residual(a, s);
    PcgScalars sc;
    RC_HIP(hipMemcpyAsync(&sc, h->scal.p, sizeof(sc), hipMemcpyDeviceToHost, s));
    std::vector<char> bt;
    if (h->nsys > 0) {
        bt.resize(batch_state_bytes(h->nsys, h->nblk));
        RC_HIP(hipMemcpyAsync(bt.data(), h->bt_state.p, bt.size(), hipMemcpyDeviceToHost, s));
    }
    RC_HIP(hipStreamSynchronize(s));
    rc_solve_stats o{};
    o.iterations = sc.iter;
    o.converged = (sc.status == 0 && sc.done && sc.iter < h->opt.maxit) ? 1 : 0;
    if (sc.status == 0 && sc.done && sc.iter == h->opt.maxit) {
        const bool met = h->opt.stop_rule == RC_STOP_REFERENCE ? (std::sqrt(sc.rz) / std::sqrt(sc.rz0) <= tol)
                                                                : (std::sqrt(sc.rr) / std::sqrt(sc.bb) <= tol);
        o.converged = met ? 1 : 0;
    }
    o.rel_res_true = sc.bnorm2 > 0.0 ? std::sqrt(sc.res2) / std::sqrt(sc.bnorm2) : 0.0;
    o.rel_res_precond = sc.rz0 > 0.0 ? std::sqrt(std::max(sc.rz, 0.0)) / std::sqrt(sc.rz0) : 0.0;
    o.pcg_ms = ms_pcg;
    o.setup_ms = ms_setup;
    o.err_vs_oracle = -1.0;
    bool any_brk = false;
    if (h->nsys > 0) {
        // per-system results; the summary reports the worst system
        BatchArgs hb;
        hb.nsys = h->nsys;
        hb.nblk = h->nblk;
        batch_carve(hb, bt.data());
        h->bstats.assign((size_t)h->nsys, rc_solve_stats{});
        int max_it = 0, all_conv = 1;
        double worst_true = 0.0, worst_pre = 0.0;
        for (int q = 0; q < h->nsys; ++q) {
            rc_solve_stats &b = h->bstats[(size_t)q];
            b.iterations = hb.iter[q];
            const double pre = hb.rz0[q] > 0.0 ? std::sqrt(std::max(hb.rz[q], 0.0)) / std::sqrt(hb.rz0[q]) : 0.0;
            const double tr = hb.bn2[q] > 0.0 ? std::sqrt(hb.res2[q]) / std::sqrt(hb.bn2[q]) : 0.0;
            const bool met = h->opt.stop_rule == RC_STOP_REFERENCE
                                 ? (hb.rz0[q] <= tol || pre <= tol)
                                 : (hb.bb[q] <= 0.0 || std::sqrt(hb.rr[q]) / std::sqrt(hb.bb[q]) <= tol);
            b.converged = (sc.status == 0 && !hb.brk[q] && hb.done[q] && met) ? 1 : 0;
            b.rel_res_true = tr;
            b.rel_res_precond = pre;
            b.pcg_ms = ms_pcg;
            b.setup_ms = ms_setup;
            b.err_vs_oracle = -1.0;
            any_brk = any_brk || hb.brk[q];
            max_it = std::max(max_it, b.iterations);
            all_conv &= b.converged;
            worst_true = std::max(worst_true, tr);
            worst_pre = std::max(worst_pre, pre);
        }
        o.iterations = max_it;
        o.converged = all_conv;
        o.rel_res_true = worst_true;
        o.rel_res_precond = worst_pre;
    }
    if (st) *st = o;
    *broke = (sc.status & kPcgBreakdown) || any_brk;
    if (sc.status & kPcgTimeout) return fail(h, RC_ERR_TIMEOUT, "triangular solve wait timed out");
    if (*broke) return fail(h, RC_ERR_BREAKDOWN, "PCG breakdown (p'Ap <= 0)");
    return RC_OK;
}

// Runs the PCGs of `count` handles (modes[i]) together: every member's init
// on its own stream, then graphs of two iterations of every member on the
// first handle's stream until all are done, then every member's end.
// st[i] points to member i's statistics (1 or k entries, or nullptr).
// Returns the first member's error (every member is finished regardless);
// a breakdown is reported only when nothing worse happened.
static int run_group(int count, const rc_handle_t *hs, const int *modes, double tol, rc_solve_stats *const *st)
{
    rc_handle_t h0 = hs[0];
    hipStream_t o = h0->stream;
    std::vector<MultiArgs> ma((size_t)count);
    std::vector<hipEvent_t> ev0((size_t)count, nullptr), ev1((size_t)count, nullptr);
    hipEvent_t el0 = nullptr, el1 = nullptr, join = nullptr;
    hipGraphExec_t exec = nullptr;
    bool own_exec = false;
    int *done_h = nullptr;
    int rc = RC_OK;
    auto cleanup = [&]() {
        for (auto e : ev0) if (e) (void)hipEventDestroy(e);
        for (auto e : ev1) if (e) (void)hipEventDestroy(e);
        if (el0) (void)hipEventDestroy(el0);
        if (el1) (void)hipEventDestroy(el1);
        if (join) (void)hipEventDestroy(join);
        if (own_exec && exec) (void)hipGraphExecDestroy(exec);
        if (done_h) (void)hipHostFree(done_h);
    };
    auto hip_fail = [&](rc_handle_t h, hipError_t e, const char *what) {
        cleanup();
        return fail(h, e == hipErrorOutOfMemory ? RC_ERR_NOMEM : RC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    hipError_t e = hipSuccess;
    for (int i = 0; i < count; ++i) {
        rc_handle_t h = hs[i];
        if ((e = hipEventCreate(&ev0[(size_t)i])) != hipSuccess || (e = hipEventCreate(&ev1[(size_t)i])) != hipSuccess)
            return hip_fail(h, e, "hipEventCreate");
        if ((e = hipEventRecord(ev0[(size_t)i], h->stream)) != hipSuccess) return hip_fail(h, e, "hipEventRecord");
        rc = run_begin(h, modes[i], tol, ma[(size_t)i]);
        if (rc != RC_OK) {
            if (i > 0) fail(h0, rc, std::string("group member ") + std::to_string(i) + ": " + h->err);
            cleanup();
            return rc;
        }
        if ((e = hipEventRecord(ev1[(size_t)i], h->stream)) != hipSuccess) return hip_fail(h, e, "hipEventRecord");
        if (i > 0 && (e = hipStreamWaitEvent(o, ev1[(size_t)i], 0)) != hipSuccess) return hip_fail(h0, e, "hipStreamWaitEvent");
    }
    if (count == 1) {
        rc = run_graph(h0, modes[0], ma[0], &exec);
        if (rc != RC_OK) { cleanup(); return rc; }
    } else {
        hipGraph_t g;
        if ((e = hipStreamBeginCapture(o, hipStreamCaptureModeThreadLocal)) != hipSuccess)
            return hip_fail(h0, e, "hipStreamBeginCapture");
        for (int i = 0; i < count; ++i) run_pair(hs[i], modes[i], ma[(size_t)i], o);
        if ((e = hipStreamEndCapture(o, &g)) != hipSuccess) return hip_fail(h0, e, "hipStreamEndCapture");
        e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) { exec = nullptr; return hip_fail(h0, e, "hipGraphInstantiate"); }
        own_exec = true;
    }
    if ((e = hipEventCreate(&el0)) != hipSuccess || (e = hipEventCreate(&el1)) != hipSuccess ||
        (e = hipEventCreate(&join)) != hipSuccess)
        return hip_fail(h0, e, "hipEventCreate");
    if ((e = hipHostMalloc((void **)&done_h, sizeof(int) * (size_t)count, hipHostMallocDefault)) != hipSuccess)
        return hip_fail(h0, e, "hipHostMalloc");
    if ((e = hipEventRecord(el0, o)) != hipSuccess) return hip_fail(h0, e, "hipEventRecord");
    // 2 iterations per graph launch; the done flags are read every 4 launches
    const int per_poll = 4;
    const int max_launch = (h0->opt.maxit + 1) / 2 + per_poll;
    int maxit_all = 0;
    for (int i = 0; i < count; ++i) maxit_all = std::max(maxit_all, hs[i]->opt.maxit);
    const int launches = (maxit_all + 1) / 2 + per_poll;
    (void)max_launch;
    for (int launched = 0; launched < launches;) {
        for (int t = 0; t < per_poll && launched < launches; ++t, ++launched)
            if ((e = hipGraphLaunch(exec, o)) != hipSuccess) return hip_fail(h0, e, "hipGraphLaunch");
        for (int i = 0; i < count; ++i)
            if ((e = hipMemcpyAsync(done_h + i, run_done_flag(hs[i], modes[i]), sizeof(int), hipMemcpyDeviceToHost,
                                    o)) != hipSuccess)
                return hip_fail(h0, e, "hipMemcpyAsync");
        if ((e = hipStreamSynchronize(o)) != hipSuccess) return hip_fail(h0, e, "hipStreamSynchronize");
        bool all = true;
        for (int i = 0; i < count; ++i) all = all && done_h[i];
        if (all) break;
    }
    if ((e = hipEventRecord(el1, o)) != hipSuccess) return hip_fail(h0, e, "hipEventRecord");
    if ((e = hipEventRecord(join, o)) != hipSuccess) return hip_fail(h0, e, "hipEventRecord");
    (void)hipEventSynchronize(el1);
    float ms_pcg = 0.f;
    (void)hipEventElapsedTime(&ms_pcg, el0, el1);
    int first = RC_OK;
    for (int i = 0; i < count; ++i) {
        rc_handle_t h = hs[i];
        if (i > 0 && (e = hipStreamWaitEvent(h->stream, join, 0)) != hipSuccess) return hip_fail(h, e, "hipStreamWaitEvent");
        float ms_setup = 0.f;
        (void)hipEventElapsedTime(&ms_setup, ev0[(size_t)i], ev1[(size_t)i]);
        bool broke = false;
        const int r = run_end(h, modes[i], tol, ma[(size_t)i], ms_setup, ms_pcg, st[i], &broke);
        if (r != RC_OK && (first == RC_OK || first == RC_ERR_BREAKDOWN)) {
            first = r;
            if (i > 0) fail(h0, r, std::string("group member ") + std::to_string(i) + ": " + h->err);
        }
    }
    cleanup();
    return first;
}

static int solve_resident(rc_handle_t h, double tol, rc_solve_stats *st)
{
    const int mode = kRunSingle;
    rc_solve_stats *sts[1] = {st};
    return run_group(1, &h, &mode, tol, sts);
}

static int rc_solve_multi_resident_impl(rc_handle_t h, double tol, rc_solve_stats *st)
{
    RC_CHECK_HANDLE();
    if (!(tol >= 0.0)) return fail(h, RC_ERR_ARG, "bad tolerance");
    int rc = run_state_ok(h, kRunMulti);
    if (rc != RC_OK) return rc;
    RC_HIP(hipSetDevice(h->opt.device));
    const int mode = kRunMulti;
    rc_solve_stats *sts[1] = {st};
    return run_group(1, &h, &mode, tol, sts);
}

static int rc_set_solve_options_impl(rc_handle_t h, int stop_rule, int maxit)
{
    RC_CHECK_HANDLE();
    if (stop_rule != RC_STOP_REFERENCE && stop_rule != RC_STOP_TRUE_RESIDUAL) return fail(h, RC_ERR_ARG, "unknown stop rule");
    if (maxit < 0) return fail(h, RC_ERR_ARG, "maxit must be >= 0");
    h->opt.stop_rule = stop_rule;
    if (maxit > 0) h->opt.maxit = maxit;
    return RC_OK;
}

static int rc_get_multi_stats_impl(rc_handle_t h, rc_solve_stats *st)
{
    RC_CHECK_HANDLE();
    if (!h->m_solved || h->m_stats.size() != (size_t)h->m_k) return fail(h, RC_ERR_STATE, "no multi-RHS solve");
    if (!st) return fail(h, RC_ERR_ARG, "null stats array");
    std::memcpy(st, h->m_stats.data(), h->m_stats.size() * sizeof(rc_solve_stats));
    return RC_OK;
}

static int rc_solve_group_impl(int count, const rc_handle_t *handles, double tol, rc_solve_stats *st)
{
    if (count <= 0 || !handles) return RC_ERR_ARG;
    for (int i = 0; i < count; ++i)
        if (!handles[i]) return RC_ERR_ARG;
    rc_handle_t h0 = handles[0];
    if (!(tol >= 0.0)) return fail(h0, RC_ERR_ARG, "bad tolerance");
    std::vector<int> modes((size_t)count);
    for (int i = 0; i < count; ++i) {
        rc_handle_t h = handles[i];
        for (int j = 0; j < i; ++j)
            if (handles[j] == h) return fail(h0, RC_ERR_ARG, "a handle appears twice in the group");
        if (h->opt.device != h0->opt.device) return fail(h0, RC_ERR_ARG, "group members on different devices");
        modes[(size_t)i] = h->rhs_mode == 2 ? kRunMulti : kRunSingle;
        const int rc = run_state_ok(h, modes[(size_t)i]);
        if (rc != RC_OK) {
            if (i > 0) fail(h0, rc, std::string("group member ") + std::to_string(i) + ": " + h->err);
            return rc;
        }
    }
    RC_HIP_H(h0, hipSetDevice(h0->opt.device));
    // member i's statistics: multi runs keep their per-RHS entries in the
    // handle (rc_get_multi_stats); the group reports one summary per member
    std::vector<std::vector<rc_solve_stats>> per((size_t)count);
    std::vector<rc_solve_stats *> sp((size_t)count);
    for (int i = 0; i < count; ++i) {
        per[(size_t)i].assign(modes[(size_t)i] == kRunMulti ? (size_t)handles[i]->m_k : 1, rc_solve_stats{});
        sp[(size_t)i] = per[(size_t)i].data();
    }
    const int rc = run_group(count, handles, modes.data(), tol, sp.data());
    if (st) {
        for (int i = 0; i < count; ++i) {
            const auto &v = per[(size_t)i];
            rc_solve_stats o = v[0];
            for (const auto &q : v) {
                o.iterations = std::max(o.iterations, q.iterations);
                o.converged = o.converged && q.converged;
                o.rel_res_true = std::max(o.rel_res_true, q.rel_res_true);
                o.rel_res_precond = std::max(o.rel_res_precond, q.rel_res_precond);
            }
            st[i] = o;
        }
    }
    return rc;
}

static int rc_get_solution_multi_impl(rc_handle_t h, double *X)
{
    RC_CHECK_HANDLE();
    if (!X) return fail(h, RC_ERR_ARG, "null X");
    if (!h->m_solved || h->m_k <= 0 || h->m_n != h->ns) return fail(h, RC_ERR_STATE, "no multi-RHS solve");
    RC_HIP(hipSetDevice(h->opt.device));
    const int k = h->m_k;
    int K = 1;
    while (K < k) K <<= 1;
    const size_t n = (size_t)h->ns;
    RC_HIP(h->m_X.ensure(n * (size_t)k * 8));
    launch_multi_scatter(h->m_x.as<double>(), h->idx_user.as<int>(), (int)n, k, K, h->m_X.as<double>(), h->stream);
    RC_HIP(hipMemcpyAsync(X, h->m_X.p, n * (size_t)k * 8, hipMemcpyDeviceToHost, h->stream));
    RC_HIP(hipStreamSynchronize(h->stream));
    return RC_OK;
}

static int rc_solve_multi_impl(rc_handle_t h, int k, const double *B, double *X, double tol, rc_solve_stats *st)
{
    RC_CHECK_HANDLE();
    if (!B || !X) return fail(h, RC_ERR_ARG, "bad multi-RHS arguments");
    if (!h->factored) return fail(h, RC_ERR_STATE, "rc_factor must precede rc_solve_multi");
    int rc = rc_set_rhs_multi(h, k, B);
    if (rc != RC_OK) return rc;
    rc = rc_solve_multi_resident(h, tol, st);
    if (rc != RC_OK && rc != RC_ERR_BREAKDOWN) return rc;
    const int rc2 = rc_get_solution_multi(h, X);
    return rc2 != RC_OK ? rc2 : rc;
}

// ------------------------------------------------------------------ timing
static const size_t kCalibBytes = (size_t)1 << 30;   // copy kernel: 1 GiB read + 1 GiB written

// Multi-RHS kernels: the SpMM and the two sweeps of the last multi-RHS
// system, k right-hand sides (bytes: SpMM 12 nnz + 4 (n+1) + 16 k n; sweep
// (4+s) nnz(L) + 4 (n+1) + 2 s k n).
static int time_multi_kernel(rc_handle_t h, int kernel, int reps, double *ms_per_launch, double *bytes_per_launch)
{
    if (!h->m_ran || h->m_k <= 0 || h->m_n != h->ns)
        return fail(h, RC_ERR_STATE, "a multi-RHS solve must precede timing its kernels");
    hipStream_t s = h->stream;
    MultiArgs a;
    int rc = multi_args(h, h->m_k, a);
    if (rc != RC_OK) return rc;
    MultiScalars sc;
    std::memset(&sc, 0, sizeof(sc));
    sc.maxit = 1 << 30;
    for (int q = 0; q < kMaxRhs; ++q) { sc.rz[q] = 1.0; sc.rz0[q] = 1.0; }
    RC_HIP(hipMemcpyAsync(h->m_sc.p, &sc, sizeof(sc), hipMemcpyHostToDevice, s));
    hipEvent_t e0, e1;
    RC_HIP(hipEventCreate(&e0));
    RC_HIP(hipEventCreate(&e1));
    double total = 0.0;
    const bool f32 = h->prec == 32;
    for (int t = 0; t < reps; ++t) {
        if (kernel == RC_KERNEL_MULTI_SPMM) {
            RC_HIP(hipEventRecord(e0, s));
            if (f32) launch_multi_spmm_only<float>(a, s); else launch_multi_spmm_only<double>(a, s);
            RC_HIP(hipEventRecord(e1, s));
        } else {
            if (f32) launch_multi_sentinels<float>(a, true, true, s); else launch_multi_sentinels<double>(a, true, true, s);
            launch_reset_counters(a.ctr_fwd, a.ctr_bwd, s);
            if (kernel == RC_KERNEL_MULTI_TRSV_BWD) {
                if (f32) launch_multi_fwd_only<float>(a, s); else launch_multi_fwd_only<double>(a, s);
            }
            RC_HIP(hipEventRecord(e0, s));
            if (kernel == RC_KERNEL_MULTI_TRSV_FWD) {
                if (f32) launch_multi_fwd_only<float>(a, s); else launch_multi_fwd_only<double>(a, s);
            } else {
                if (f32) launch_multi_bwd_only<float>(a, s); else launch_multi_bwd_only<double>(a, s);
            }
            RC_HIP(hipEventRecord(e1, s));
        }
        RC_HIP(hipGetLastError());
        RC_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        RC_HIP(hipEventElapsedTime(&ms, e0, e1));
        total += ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    RC_HIP(hipMemcpy(&sc, h->m_sc.p, sizeof(sc), hipMemcpyDeviceToHost));
    h->m_solved = false;             // the timing overwrote the solve's vectors
    if (sc.status & kPcgTimeout) return fail(h, RC_ERR_TIMEOUT, "timed kernel wait expired");
    *ms_per_launch = total / reps;
    if (bytes_per_launch) {
        const double n = (double)h->ns, k = (double)h->m_k, sz = h->prec == 32 ? 4.0 : 8.0;
        if (kernel == RC_KERNEL_MULTI_SPMM) *bytes_per_launch = 12.0 * (double)h->nnz_sys + 4.0 * (n + 1) + 16.0 * k * n;
        else *bytes_per_launch = (4.0 + sz) * (double)h->nnz_lt + 4.0 * (n + 1) + 2.0 * sz * k * n;
    }
    return RC_OK;
}

static int rc_time_kernel_impl(rc_handle_t h, int kernel, int reps, double *ms_per_launch, double *bytes_per_launch)
{
    RC_CHECK_HANDLE();
    if (h->host_only) return fail(h, RC_ERR_NO_DEVICE, "host-only handle");
    if (kernel == RC_KERNEL_MULTI_SPMM || kernel == RC_KERNEL_MULTI_TRSV_FWD || kernel == RC_KERNEL_MULTI_TRSV_BWD) {
        if (reps <= 0 || !ms_per_launch) return fail(h, RC_ERR_ARG, "bad timing arguments");
        RC_HIP(hipSetDevice(h->opt.device));
        return time_multi_kernel(h, kernel, reps, ms_per_launch, bytes_per_launch);
    }
    if (!h->factored || !h->rhs_set) return fail(h, RC_ERR_STATE, "factor and solve once before timing");
    {
        const int rc_alloc = alloc_solve(h);     // sizes the vectors, zeroes their dummy slot
        if (rc_alloc != RC_OK) return rc_alloc;
    }
    if (reps <= 0 || !ms_per_launch) return fail(h, RC_ERR_ARG, "bad timing arguments");
    RC_HIP(hipSetDevice(h->opt.device));
    hipStream_t s = h->stream;
    SolveArgs a = solve_args(h);
    PcgScalars sc;
    std::memset(&sc, 0, sizeof(sc));
    sc.tol = 0.0;
    sc.maxit = 1 << 30;
    sc.alpha = 0.0;
    sc.rz = 1.0;
    sc.rz0 = 1.0;
    RC_HIP(hipMemcpyAsync(h->scal.p, &sc, sizeof(sc), hipMemcpyHostToDevice, s));
    const double n = (double)h->ns, nnzA = (double)h->nnz_sys, nnzL = (double)h->nnz_lt;
    const double sz = h->prec == 32 ? 4.0 : 8.0;
    const double b_spmv = 12.0 * nnzA + 4.0 * (n + 1) + 16.0 * n;
    const double b_tri = (4.0 + sz) * nnzL + 4.0 * (n + 1) + 2.0 * sz * n;
    hipEvent_t e0, e1;
    RC_HIP(hipEventCreate(&e0));
    RC_HIP(hipEventCreate(&e1));
    double total = 0.0;
    auto fill_y = [&]() {
        if (h->prec == 32) launch_fill_sentinel<float>((float *)h->y.p, h->ns, s);
        else launch_fill_sentinel<double>((double *)h->y.p, h->ns, s);
    };
    auto fill_z = [&]() {
        if (h->prec == 32) launch_fill_sentinel<float>((float *)h->z.p, h->ns, s);
        else launch_fill_sentinel<double>((double *)h->z.p, h->ns, s);
    };
    for (int t = 0; t < reps; ++t) {
        float ms = 0.f;
        if (kernel == RC_KERNEL_SPMV) {
            RC_HIP(hipEventRecord(e0, s));
            if (h->prec == 32) launch_spmv_only<float>(a, s); else launch_spmv_only<double>(a, s);
            RC_HIP(hipEventRecord(e1, s));
        } else if (kernel == RC_KERNEL_TRSV_FWD) {
            fill_y();
            launch_reset_counters(a.ctr_fwd, a.ctr_bwd, s);
            RC_HIP(hipEventRecord(e0, s));
            if (h->prec == 32) launch_trsv_fwd_only<float>(a, s); else launch_trsv_fwd_only<double>(a, s);
            RC_HIP(hipEventRecord(e1, s));
        } else if (kernel == RC_KERNEL_TRSV_BWD) {
            fill_y();
            fill_z();
            launch_reset_counters(a.ctr_fwd, a.ctr_bwd, s);
            if (h->prec == 32) launch_trsv_fwd_only<float>(a, s); else launch_trsv_fwd_only<double>(a, s);
            RC_HIP(hipEventRecord(e0, s));
            if (h->prec == 32) launch_trsv_bwd_only<float>(a, s); else launch_trsv_bwd_only<double>(a, s);
            RC_HIP(hipEventRecord(e1, s));
        } else if (kernel == RC_KERNEL_PCG_ITER) {
            RC_HIP(hipEventRecord(e0, s));
            if (h->prec == 32) launch_pcg_iteration<float>(a, t & 1, s); else launch_pcg_iteration<double>(a, t & 1, s);
            RC_HIP(hipEventRecord(e1, s));
        } else if (kernel == RC_KERNEL_COPY) {
            RC_HIP(h->calib.ensure(2 * kCalibBytes));
            double *src = h->calib.as<double>();
            if (t == 0) RC_HIP(hipMemsetAsync(src, 0, 2 * kCalibBytes, s));
            RC_HIP(hipEventRecord(e0, s));
            launch_copy(src, src + kCalibBytes / 8, kCalibBytes / 8, s);
            RC_HIP(hipEventRecord(e1, s));
        } else {
            return fail(h, RC_ERR_ARG, "unknown kernel id");
        }
        RC_HIP(hipGetLastError());
        RC_HIP(hipEventSynchronize(e1));
        RC_HIP(hipEventElapsedTime(&ms, e0, e1));
        total += ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    unsigned int stat = 0;
    RC_HIP(hipMemcpy(&stat, (char *)h->scal.p + offsetof(PcgScalars, status), sizeof(stat), hipMemcpyDeviceToHost));
    if (stat & kPcgTimeout) return fail(h, RC_ERR_TIMEOUT, "timed kernel wait expired");
    *ms_per_launch = total / reps;
    if (bytes_per_launch) {
        if (kernel == RC_KERNEL_SPMV) *bytes_per_launch = b_spmv;
        else if (kernel == RC_KERNEL_COPY) *bytes_per_launch = 2.0 * (double)kCalibBytes;
        else if (kernel == RC_KERNEL_PCG_ITER) *bytes_per_launch = b_spmv + 2.0 * b_tri + sz * n + 72.0 * n;
        else *bytes_per_launch = b_tri;
    }
    return RC_OK;
}

// Resident-input variants used by the benchmark: the RHS is uploaded and
// permuted before the timed region, the solution stays on the device.
static int rc_set_rhs_impl(rc_handle_t h, const double *b)
{
    RC_CHECK_HANDLE();
    if (h->host_only) return fail(h, RC_ERR_NO_DEVICE, "host-only handle");
    if (!b) return fail(h, RC_ERR_ARG, "null rhs");
    if (!h->factored) return fail(h, RC_ERR_STATE, "rc_factor must precede rc_set_rhs");
    RC_HIP(hipSetDevice(h->opt.device));
    int rc = set_rhs_device(h, b);
    if (rc != RC_OK) return rc;
    RC_HIP(hipStreamSynchronize(h->stream));
    return RC_OK;
}

static int rc_solve_resident_impl(rc_handle_t h, double tol, rc_solve_stats *st)
{
    RC_CHECK_HANDLE();
    if (h->host_only) return fail(h, RC_ERR_NO_DEVICE, "host-only handle");
    if (!h->factored || !h->rhs_set) return fail(h, RC_ERR_STATE, "factor and rc_set_rhs must precede");
    RC_HIP(hipSetDevice(h->opt.device));
    return solve_resident(h, tol, st);
}

// Autotuner: per factored structure, time both sweeps (forward + backward,
// hipEvents, `reps` launches each) for every candidate (tail width limit x
// sweep grid) and keep the fastest.  Changing the tail limit re-runs only the
// assembly (the elimination's output stays on the device).
static int rc_autotune_impl(rc_handle_t h, int reps, rc_tune_stats *st)
{
    RC_CHECK_HANDLE();
    if (h->host_only) return fail(h, RC_ERR_NO_DEVICE, "host-only handle");
    if (!h->factored || !h->rhs_set) return fail(h, RC_ERR_STATE, "factor and rc_set_rhs must precede rc_autotune");
    if (reps <= 0) reps = 5;
    RC_HIP(hipSetDevice(h->opt.device));
    const int tails[] = {0, 64, 128, 256, kTailMaxRows + 1};
    const int grids[] = {h->sweep_blocks_max, h->sweep_blocks_max / 2, h->sweep_blocks_max / 4};
    auto reassemble = [&](int tail) -> int {
        h->tail_max_rows = tail;
        double ms = 0.0;
        int rc = (h->prec == 32) ? assemble<float>(h, &ms) : assemble<double>(h, &ms);
        invalidate_graph(h);
        return rc;
    };
    auto sweep_ms = [&](double *out) -> int {
        double f = 0.0, b = 0.0;
        int rc = rc_time_kernel(h, RC_KERNEL_TRSV_FWD, reps, &f, nullptr);
        if (rc == RC_OK) rc = rc_time_kernel(h, RC_KERNEL_TRSV_BWD, reps, &b, nullptr);
        *out = f + b;
        return rc;
    };
    const int tail0 = h->tail_max_rows, grid0 = h->sweep_blocks;
    double base = 0.0;
    int rc = sweep_ms(&base);
    if (rc != RC_OK) return rc;
    double best = base;
    int best_tail = tail0, best_grid = grid0, tried = 1;
    for (int tail : tails) {
        rc = reassemble(tail);
        if (rc != RC_OK) return rc;
        for (int g : grids) {
            if (g < 1) continue;
            h->sweep_blocks = g;
            double ms = 0.0;
            rc = sweep_ms(&ms);
            if (rc != RC_OK) return rc;
            ++tried;
            if (ms < best) { best = ms; best_tail = tail; best_grid = g; }
        }
    }
    h->sweep_blocks = best_grid;
    rc = reassemble(best_tail);
    if (rc != RC_OK) return rc;
    h->fstats.nlevels = h->nlevels;
    if (st) {
        st->tail_max_rows = best_tail;
        st->sweep_blocks = best_grid;
        st->configs_tried = tried;
        st->sweeps_ms_before = base;
        st->sweeps_ms_after = best;
    }
    return RC_OK;
}

static int rc_get_solution_impl(rc_handle_t h, double *x)
{
    RC_CHECK_HANDLE();
    if (!x) return fail(h, RC_ERR_ARG, "null x");
    if (!h->rhs_set) return fail(h, RC_ERR_STATE, "no solve");
    return fetch_solution(h, x);
}

}  // extern "C"

extern "C" {

int rc_create(rc_handle_t *out, const rc_options *opt)
{
    return guarded(nullptr, [&]() { return rc_create_impl(out, opt); });
}

int rc_destroy(rc_handle_t h)
{
    return guarded(h, [&]() { return rc_destroy_impl(h); });
}

int rc_set_triplets(rc_handle_t h, int n, int64_t m, const int *rows, const int *cols, const double *vals, int kind)
{
    return guarded(h, [&]() { return rc_set_triplets_impl(h, n, m, rows, cols, vals, kind); });
}

int rc_set_batch(rc_handle_t h, int nsys, const int *sys_n, const int64_t *sys_m, const int *rows, const int *cols, const double *vals)
{
    return guarded(h, [&]() { return rc_set_batch_impl(h, nsys, sys_n, sys_m, rows, cols, vals); });
}

int rc_get_batch(rc_handle_t h, int *nsys, int64_t *offsets)
{
    return guarded(h, [&]() { return rc_get_batch_impl(h, nsys, offsets); });
}

int rc_load_mtx(rc_handle_t h, const char *path, int kind)
{
    return guarded(h, [&]() { return rc_load_mtx_impl(h, path, kind); });
}

int rc_get_size(rc_handle_t h, int64_t *n, int64_t *nnz)
{
    return guarded(h, [&]() { return rc_get_size_impl(h, n, nnz); });
}

int rc_set_locality(rc_handle_t h, int mode)
{
    return guarded(h, [&]() { return rc_set_locality_impl(h, mode); });
}

int rc_get_locality(rc_handle_t h, int *rank, double *seconds)
{
    return guarded(h, [&]() { return rc_get_locality_impl(h, rank, seconds); });
}

int rc_order(rc_handle_t h, int method, uint64_t seed)
{
    return guarded(h, [&]() { return rc_order_impl(h, method, seed); });
}

int rc_get_permutation(rc_handle_t h, int *perm)
{
    return guarded(h, [&]() { return rc_get_permutation_impl(h, perm); });
}

int rc_set_gnn_weights(rc_handle_t h, const float *w, int64_t count)
{
    return guarded(h, [&]() { return rc_set_gnn_weights_impl(h, w, count); });
}

int rc_set_permutation(rc_handle_t h, const int *perm)
{
    return guarded(h, [&]() { return rc_set_permutation_impl(h, perm); });
}

int rc_gnn_scores(rc_handle_t h, const float *w, float *scores)
{
    return guarded(h, [&]() { return rc_gnn_scores_impl(h, w, scores); });
}

int rc_gnn_gradient(rc_handle_t h, const float *w, const float *dscores, float *dw)
{
    return guarded(h, [&]() { return rc_gnn_gradient_impl(h, w, dscores, dw); });
}

int rc_get_prepared_csc(rc_handle_t h, int64_t *n, int64_t *nnz, int64_t *colptr, int *rowind, double *val)
{
    return guarded(h, [&]() { return rc_get_prepared_csc_impl(h, n, nnz, colptr, rowind, val); });
}

int rc_write_prepared_mtx(rc_handle_t h, const char *path)
{
    return guarded(h, [&]() { return rc_write_prepared_mtx_impl(h, path); });
}

int rc_factor(rc_handle_t h, rc_factor_stats *st)
{
    return guarded(h, [&]() { return rc_factor_impl(h, st); });
}

int rc_get_factor(rc_handle_t h, int64_t *n, int64_t *nnz, int64_t *rowptr, int *colind, double *val, double *diag)
{
    return guarded(h, [&]() { return rc_get_factor_impl(h, n, nnz, rowptr, colind, val, diag); });
}

int rc_dump_factor(rc_handle_t h, const char *prefix)
{
    return guarded(h, [&]() { return rc_dump_factor_impl(h, prefix); });
}

int rc_load_factor(rc_handle_t h, const char *prefix)
{
    return guarded(h, [&]() { return rc_load_factor_impl(h, prefix); });
}

int rc_make_rhs(rc_handle_t h, int kind, uint64_t seed, double *b)
{
    return guarded(h, [&]() { return rc_make_rhs_impl(h, kind, seed, b); });
}

int rc_set_reference_solution(rc_handle_t h, const double *x_ref)
{
    return guarded(h, [&]() { return rc_set_reference_solution_impl(h, x_ref); });
}

int rc_solve(rc_handle_t h, const double *b, double *x, double tol, rc_solve_stats *st)
{
    return guarded(h, [&]() { return rc_solve_impl(h, b, x, tol, st); });
}

int rc_solve_batch(rc_handle_t h, const double *b, double *x, double tol, rc_solve_stats *per_system)
{
    return guarded(h, [&]() { return rc_solve_batch_impl(h, b, x, tol, per_system); });
}

int rc_get_batch_stats(rc_handle_t h, rc_solve_stats *per_system)
{
    return guarded(h, [&]() { return rc_get_batch_stats_impl(h, per_system); });
}

int rc_set_rhs_multi(rc_handle_t h, int k, const double *B)
{
    return guarded(h, [&]() { return rc_set_rhs_multi_impl(h, k, B); });
}

int rc_solve_multi_resident(rc_handle_t h, double tol, rc_solve_stats *st)
{
    return guarded(h, [&]() { return rc_solve_multi_resident_impl(h, tol, st); });
}

int rc_get_solution_multi(rc_handle_t h, double *X)
{
    return guarded(h, [&]() { return rc_get_solution_multi_impl(h, X); });
}

int rc_solve_multi(rc_handle_t h, int k, const double *B, double *X, double tol, rc_solve_stats *st)
{
    return guarded(h, [&]() { return rc_solve_multi_impl(h, k, B, X, tol, st); });
}

int rc_time_kernel(rc_handle_t h, int kernel, int reps, double *ms_per_launch, double *bytes_per_launch)
{
    return guarded(h, [&]() { return rc_time_kernel_impl(h, kernel, reps, ms_per_launch, bytes_per_launch); });
}

int rc_set_rhs(rc_handle_t h, const double *b)
{
    return guarded(h, [&]() { return rc_set_rhs_impl(h, b); });
}

int rc_solve_resident(rc_handle_t h, double tol, rc_solve_stats *st)
{
    return guarded(h, [&]() { return rc_solve_resident_impl(h, tol, st); });
}

int rc_autotune(rc_handle_t h, int reps, rc_tune_stats *st)
{
    return guarded(h, [&]() { return rc_autotune_impl(h, reps, st); });
}

int rc_get_solution(rc_handle_t h, double *x)
{
    return guarded(h, [&]() { return rc_get_solution_impl(h, x); });
}

int rc_factor_ic0(rc_handle_t h, rc_factor_stats *st)
{
    return guarded(h, [&]() { return rc_factor_ic0_impl(h, st); });
}

int rc_set_solve_options(rc_handle_t h, int stop_rule, int maxit)
{
    return guarded(h, [&]() { return rc_set_solve_options_impl(h, stop_rule, maxit); });
}

int rc_get_multi_stats(rc_handle_t h, rc_solve_stats *st)
{
    return guarded(h, [&]() { return rc_get_multi_stats_impl(h, st); });
}

int rc_solve_group(int count, const rc_handle_t *handles, double tol, rc_solve_stats *st)
{
    rc_handle_t h0 = (count > 0 && handles) ? handles[0] : nullptr;
    return guarded(h0, [&]() { return rc_solve_group_impl(count, handles, tol, st); });
}

}  // extern "C"
