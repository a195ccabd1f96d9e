// pcg.hip -- PCG with the randomized-Cholesky preconditioner on gfx950.
//
// Follows gpu_implementation/solver.hpp:149-233 (z = L^-T D+ L^-1 r, D+
// divides only where D != 0) and solver.hpp:457-569 (x0 = 0; alpha =
// r'z / p'Ap; x += alpha p; r -= alpha Ap; z = M r; beta = r'z_new / r'z;
// p = z + beta p; stop on sqrt(r'z / r0'z0) <= tol or k = maxit).
//
// One iteration is a handful of launches with device-resident scalars and no
// host synchronisation, so a run of iterations is captured in one hipGraph:
//   K1 p_new = z + beta p_old (streaming), then the SpMV Ap = A p_new with
//      block partials of p'Ap; the last block computes alpha;
//   K2 fwd    r -= alpha Ap and x += alpha p on the row being solved, then the
//             forward sweep y = L^-1 r: persistent chunks over the wide levels,
//             then the narrow trailing levels in one LDS-resident workgroup;
//   K3 bwd    z = L^-T (D+ y): the trailing levels first, then the wide ones;
//   K4 dot    r'z block partials, last block computes beta and the stop test.
// All vectors live in the level-order position numbering, so a chunk of the
// sweeps (and a 64-row tile of the SpMV) reads and writes its own rows with
// coalesced accesses; L, L^T and A are stored as sliced ELL, so index/value
// loads are coalesced too.  Padding entries point at the dummy slot n, which
// always holds zero.  In the persistent sweeps a workgroup claims the next
// chunk of the level order; a row issues its dependency loads eight at a time
// with plain (cached) loads and, only for a value still holding the sentinel,
// polls it with agent-scope loads.  Values are written once per sweep with
// agent-scope stores, so no fences are needed; each consumer resets the vector
// it drains back to the sentinel.  All reductions use a fixed grid and a fixed
// order.
#include "common.cuh"
#include "kernels.h"

namespace rc {

namespace {

constexpr int kRedBlock = 256;

__device__ __forceinline__ double block_sum(double v, double *lds)
{
    v = wave_sum(v);
    const int wid = threadIdx.x / kWave;
    if (lane_id() == 0) lds[wid] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += lds[w];
    __syncthreads();
    return t;   // valid in thread 0
}

// Sum of partials[0..m) by one block in a fixed order; valid in thread 0.
__device__ __forceinline__ double reduce_partials(const double *partials, int m, double *lds)
{
    double v = 0.0;
    for (int i = threadIdx.x; i < m; i += blockDim.x) v += ld_relaxed_val(partials + i);
    return block_sum(v, lds);
}

// Returns true in every thread of the block that arrived last.
__device__ __forceinline__ bool last_block(unsigned int *arrive, int *lds_flag)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        fence_release();
        unsigned int old = atomicAdd(arrive, 1u);
        *lds_flag = (old == gridDim.x - 1) ? 1 : 0;
        if (*lds_flag) fence_acquire();
    }
    __syncthreads();
    return *lds_flag != 0;
}

template <typename T>
__device__ __noinline__ T wait_slow(T *p, long long spin_ticks, PcgScalars *sc, int nap_max)
{
    long long t0 = wall_clock64();
    int nap = 1;
    while (true) {
        T v = ld_relaxed_val(p);
        if (!is_sent(v)) return v;
        if (ld_relaxed(&sc->status) & kPcgTimeout) return (T)0;
        if (wall_clock64() - t0 > spin_ticks) {
            atomicOr(&sc->status, kPcgTimeout);
            return (T)0;
        }
        // exponential back-off keeps waiting waves from flooding memory
        for (int q = 0; q < nap; ++q) __builtin_amdgcn_s_sleep(2);
        nap = nap < nap_max ? nap * 2 : nap_max;
    }
}

// acc - sum_s val[s] * vec[ind[s]] over this lane's slices of a grouped
// sliced-ELL row (slice stride `stride`), in slice order.  Dependency loads
// are issued eight at a time; a value still holding the sentinel is polled.
template <typename T>
__device__ __forceinline__ T grouped_row_solve(const int *ind, const T *val, int nsl, int stride, T *vec, T acc,
                                               long long spin, PcgScalars *sc, int nap_max)
{
    constexpr int B = 8;
    int sl = 0;
    for (; sl + B <= nsl; sl += B) {
        int j[B];
        T l[B], v[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            j[u] = ind[(sl + u) * stride];
            l[u] = val[(sl + u) * stride];
        }
#pragma unroll
        for (int u = 0; u < B; ++u) v[u] = vec[j[u]];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            if (is_sent(v[u])) v[u] = wait_slow(vec + j[u], spin, sc, nap_max);
            acc = fma_t(-l[u], v[u], acc);
        }
    }
    for (; sl < nsl; ++sl) {
        int j = ind[sl * stride];
        T l = val[sl * stride];
        T v = vec[j];
        if (is_sent(v)) v = wait_slow(vec + j, spin, sc, nap_max);
        acc = fma_t(-l, v, acc);
    }
    return acc;
}

// The same sum when the row's first kBulkPref slices are already in
// registers (jp/lp, loaded a chunk earlier; a lane without slice u holds the
// zero entry, whose value is 0 and whose index is the dummy slot).  Slices
// past them are loaded here as above.
constexpr int kBulkPref = 8;
template <typename T>
__device__ __forceinline__ T grouped_row_solve_pref(const int *jp, const T *lp, const int *ind, const T *val, int nsl,
                                                    int stride, T *vec, T acc, long long spin, PcgScalars *sc,
                                                    int nap_max)
{
    T v[kBulkPref];
#pragma unroll
    for (int u = 0; u < kBulkPref; ++u) v[u] = vec[jp[u]];
#pragma unroll
    for (int u = 0; u < kBulkPref; ++u) {
        if (is_sent(v[u])) v[u] = wait_slow(vec + jp[u], spin, sc, nap_max);
        acc = fma_t(-lp[u], v[u], acc);
    }
    if (nsl > kBulkPref)
        acc = grouped_row_solve<T>(ind + kBulkPref * stride, val + kBulkPref * stride, nsl - kBulkPref, stride, vec,
                                   acc, spin, sc, nap_max);
    return acc;
}

// Sum over the G lanes of a row group (G a power of two <= 64, groups aligned
// inside the wave), valid in the group's LAST lane: a shift-and-add scan over
// DPP row shifts (1, 2, 4, 8 lanes inside 16-lane rows) and row broadcasts
// (lane 15 to the next row, lane 31 to the upper half).  The lanes move data
// through the VALU, not through LDS like a shuffle; the order is fixed, so the
// result is deterministic.  Every lane of the wave must be active.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ float dpp_move(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, kRowMask, 0xF, true));
}
template <int kCtrl, int kRowMask>
__device__ __forceinline__ double dpp_move(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), kCtrl, kRowMask, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), kCtrl, kRowMask, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <typename T>
__device__ __forceinline__ T group_sum(T v, int G)
{
    if (G >= 2) v += dpp_move<0x111, 0xF>(v);     // row_shr:1
    if (G >= 4) v += dpp_move<0x112, 0xF>(v);     // row_shr:2
    if (G >= 8) v += dpp_move<0x114, 0xF>(v);     // row_shr:4
    if (G >= 16) v += dpp_move<0x118, 0xF>(v);    // row_shr:8
    if (G >= 32) v += dpp_move<0x142, 0xA>(v);    // row_bcast:15 into rows 1 and 3
    if (G >= 64) v += dpp_move<0x143, 0xC>(v);    // row_bcast:31 into rows 2 and 3
    return v;
}

// Thread range of a grid-stride loop over the ordered SpMV tiles in which
// the workgroups of one XCD (workgroup b runs on XCD b % 8) share one
// contiguous eighth of the tile order, so each XCD's L2 serves one region of
// the input numbering.  Falls back to the plain grid stride when the grid is
// not a multiple of 8.
struct TileRange {
    int first, end, stride;
};
__device__ __forceinline__ TileRange xcd_tile_range(int ntiles)
{
    constexpr int kXcds = 8;
    TileRange r;
    if (gridDim.x % kXcds != 0) {
        r.first = blockIdx.x * blockDim.x + threadIdx.x;
        r.end = ntiles * kWave;
        r.stride = gridDim.x * blockDim.x;
        return r;
    }
    const int x = blockIdx.x % kXcds, b = blockIdx.x / kXcds;
    const int per = (ntiles + kXcds - 1) / kXcds;
    const int t0 = min(ntiles, x * per), t1 = min(ntiles, t0 + per);
    r.first = t0 * kWave + b * blockDim.x + threadIdx.x;
    r.end = t1 * kWave;
    r.stride = (gridDim.x / kXcds) * blockDim.x;
    return r;
}

// ---------------------------------------------------------------- K1
// p_new = z + beta p_old (streaming), then Ap = A p_new with block partials of
// p'Ap; the last block computes alpha.  (Forming p_new inside the SpMV gather
// would recompute it once per matrix entry.)
template <typename T>
__global__ void __launch_bounds__(kRedBlock) k_pupdate(SolveArgs a, const double *p_old, double *p_new)
{
    PcgScalars *sc = a.sc;
    if (sc->done) return;
    const double beta = sc->beta;
    const T *z = (const T *)a.z;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x)
        p_new[i] = fma_t(beta, p_old[i], (double)z[i]);
}

template <typename T>
__global__ void __launch_bounds__(kRedBlock) k_spmv(SolveArgs a, const double *p)
{
    __shared__ double lds[kRedBlock / kWave];
    __shared__ int flag;
    PcgScalars *sc = a.sc;
    if (sc->done) return;
    if (blockIdx.x == 0) reset_claims(a.ctr_fwd, a.ctr_bwd, threadIdx.x);
    const int lane = lane_id();
    double acc_dot = 0.0;
    const TileRange tr = xcd_tile_range(a.ntiles);
    for (int it = tr.first; it < tr.end; it += tr.stride) {
        const int tile = a.a_tile_ord[it / kWave];
        const int q = tile * kWave + lane;
        const int w = a.a_width[tile];
        const int *ind = a.a_ind + a.a_off[tile] + lane;
        const double *val = a.a_val + a.a_off[tile] + lane;
        double acc = 0.0;
        int t = 0;
        for (; t + 4 <= w; t += 4) {
            int j[4];
            double v[4], pj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) { j[u] = ind[(t + u) * kWave]; v[u] = val[(t + u) * kWave]; }
#pragma unroll
            for (int u = 0; u < 4; ++u) pj[u] = p[j[u]];
#pragma unroll
            for (int u = 0; u < 4; ++u) acc = fma_t(v[u], pj[u], acc);
        }
        for (; t < w; ++t) acc = fma_t(val[t * kWave], p[ind[t * kWave]], acc);
        if (q < a.n) {
            a.ap[q] = acc;
            acc_dot += p[q] * acc;
        }
    }
    double bs = block_sum(acc_dot, lds);
    if (threadIdx.x == 0) st_relaxed_val(a.partials + blockIdx.x, bs);
    if (last_block(&sc->arrive_spmv, &flag)) {
        double pap = reduce_partials(a.partials, gridDim.x, lds);
        if (threadIdx.x == 0) {
            sc->pap = pap;
            if (!(pap > 0.0) || !isfinite(pap)) {      // breakdown (solver.hpp has no check)
                sc->status |= kPcgBreakdown;
                sc->done = 1;
                sc->alpha = 0.0;
            } else {
                sc->alpha = sc->rz / pap;
            }
            sc->arrive_spmv = 0u;
        }
    }
}

// Plain SpMV for timing: y = A x over the same sliced-ELL layout.
__global__ void __launch_bounds__(kRedBlock) k_spmv_plain(SolveArgs a, const double *x, double *y)
{
    const int lane = lane_id();
    const TileRange tr = xcd_tile_range(a.ntiles);
    for (int it = tr.first; it < tr.end; it += tr.stride) {
        const int tile = a.a_tile_ord[it / kWave];
        const int p = tile * kWave + lane;
        const int w = a.a_width[tile];
        const int *ind = a.a_ind + a.a_off[tile] + lane;
        const double *val = a.a_val + a.a_off[tile] + lane;
        double acc = 0.0;
        int t = 0;
        for (; t + 4 <= w; t += 4) {
            int j[4];
            double v[4], xj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) { j[u] = ind[(t + u) * kWave]; v[u] = val[(t + u) * kWave]; }
#pragma unroll
            for (int u = 0; u < 4; ++u) xj[u] = x[j[u]];
#pragma unroll
            for (int u = 0; u < 4; ++u) acc = fma_t(v[u], xj[u], acc);
        }
        for (; t < w; ++t) acc = fma_t(val[t * kWave], x[ind[t * kWave]], acc);
        if (p < a.n) y[p] = acc;
    }
}

// ---------------------------------------------------------------- K2/K3
// Chunk m = {start, cnt, ELL offset, slices} of the forward sweep, processed
// by one 256-thread group (thread t of the group); G lanes per row.  The row's
// own values and the chunk's index/value slices are loaded together, then
// the dependencies.
template <typename T>
__device__ __forceinline__ void fwd_chunk(const SolveArgs &a, int4 m, int t, double alpha, const double *p_cur,
                                          int update_rx, const int *jp, const T *lp)
{
    const T *lval = (const T *)a.lf_val;
    T *y = (T *)a.y;
    T *z = (T *)a.z;
    const int cnt = m.y;
    const int G = sweep_group(cnt);
    const int stride = sweep_rows_pow2(cnt) * G;
    const int i = t / G, g = t - i * G;
    const bool active = t < stride && i < cnt;
    const bool lead = active && g == G - 1;     // the lane that receives the group sum
    const int k = m.x + i;
    T acc = (T)0;
    double rk = 0.0, apk = 0.0, xk = 0.0, pk = 0.0;
    if (lead) {
        rk = a.r[k];
        if (update_rx) {
            apk = a.ap[k];
            xk = a.x[k];
            pk = p_cur[k];
            if (a.alpha_sys) alpha = a.alpha_sys[a.sys_of[k]];
        }
    }
    if (t < stride) {
        const int off = m.z + t;
        acc = grouped_row_solve_pref<T>(jp, lp, a.lf_ind + off, lval + off, m.w, stride, y, acc, a.spin_ticks, a.sc,
                                        a.nap_max);
    }
    if (lead) {
        if (update_rx) {
            rk = fma_t(-alpha, apk, rk);
            a.r[k] = rk;
            a.x[k] = fma_t(alpha, pk, xk);
        }
        acc += (T)rk;
    }
    acc = group_sum(acc, G);
    if (lead) {
        st_relaxed_val(y + k, unsent(acc));
        z[k] = sent_value(T());
    }
}

template <typename T>
__device__ __forceinline__ void bwd_chunk(const SolveArgs &a, int4 m, int t, const int *jp, const T *lp)
{
    const T *ltval = (const T *)a.lb_val;
    const T *diag = (const T *)a.diag;
    T *y = (T *)a.y;
    T *z = (T *)a.z;
    const int cnt = m.y;
    const int G = sweep_group(cnt);
    const int stride = sweep_rows_pow2(cnt) * G;
    const int i = t / G, g = t - i * G;
    const bool active = t < stride && i < cnt;
    const bool lead = active && g == G - 1;     // the lane that receives the group sum
    const int k = m.x + i;
    T acc = (T)0, yk = (T)0, dk = (T)0;
    if (lead) {
        yk = y[k];
        dk = diag[k];
    }
    if (t < stride) {
        const int off = m.z + t;
        acc = grouped_row_solve_pref<T>(jp, lp, a.lb_ind + off, ltval + off, m.w, stride, z, acc, a.spin_ticks, a.sc,
                                        a.nap_max);
    }
    if (lead) {
        y[k] = sent_value(T());
        if (dk != (T)0) yk = yk / dk;            // solver.hpp:64-84
        acc += yk;
    }
    acc = group_sum(acc, G);
    if (lead) st_relaxed_val(z + k, unsent(acc));
}

// Persistent part: workgroups claim chunks of the wide levels in order
// (forward) or reverse order (backward); dependencies are awaited per value.
// Claims run ahead of the work: while chunk i is processed, the workgroup
// already holds chunk i+1's descriptor and its first kBulkPref index/value
// slices (loaded before chunk i's dependency gathers, so both round trips
// overlap), loads chunk i+2's descriptor, and thread 0 has the claim of chunk
// i+3 in flight, so no round trip is exposed between chunks.  Holding claims
// ahead cannot deadlock: the smallest unfinished claimed chunk is always some
// workgroup's current chunk, and its dependencies are all earlier chunks.
template <typename T, bool kFwd>
__global__ void __launch_bounds__(kSweepThreads) k_sweep_bulk(SolveArgs a, const double *p_cur, int update_rx)
{
    __shared__ int s_init[3];
    __shared__ int s_claim[2];
    __shared__ int s_stop;
    PcgScalars *sc = a.sc;
    if (sc->done) return;
    const double alpha = (kFwd && update_rx) ? sc->alpha : 0.0;
    // claim queue of this workgroup: its j-th claim is chunk j * nq + q
    const int nq = (int)gridDim.x < kClaimQueues ? (int)gridDim.x : kClaimQueues;
    const int q = (int)(blockIdx.x % (unsigned int)nq);
    unsigned int *ctr = (kFwd ? a.ctr_fwd : a.ctr_bwd) + q * kClaimStride;
    const int4 *meta = kFwd ? a.cm_f : a.cm_b;
    const int *ind = kFwd ? a.lf_ind : a.lb_ind;
    const T *val = (const T *)(kFwd ? a.lf_val : a.lb_val);
    const int ze = kFwd ? a.lf_zero : a.lb_zero;
    const int nb = a.bulk_chunks;
    const int t = threadIdx.x;
    // descriptor of claim c (clamped: a claim past the end is never processed).
    // The lane term (a.zero is 0) keeps the load a per-lane one, so the
    // compiler does not move the value to scalar registers (and wait for it)
    // at the load; it is made uniform only when it is used, a chunk later.
    auto desc = [&](int c) {
        const int cc = c < nb ? c : nb - 1;
        return meta[(kFwd ? cc : nb - 1 - cc) + lane_id() * a.zero];
    };
    auto uniform = [](int4 v) {
        return make_int4(__builtin_amdgcn_readfirstlane(v.x), __builtin_amdgcn_readfirstlane(v.y),
                         __builtin_amdgcn_readfirstlane(v.z), __builtin_amdgcn_readfirstlane(v.w));
    };
    // this lane's first kBulkPref slices of chunk m; every lane loads (the
    // zero entry where it has no slice), and nothing is computed from the
    // values until the chunk is processed
    auto fetch = [&](const int4 &m, int *j, T *l) {
        const int stride = sweep_rows_pow2(m.y) * sweep_group(m.y);
#pragma unroll
        for (int u = 0; u < kBulkPref; ++u) {
            const bool has = t < stride && u < m.w;
            const int e = has ? m.z + t + u * stride : ze;
            j[u] = ind[e];
            l[u] = val[e];
        }
    };
    int c_pend = 0;
    unsigned int st_pend = 0;
    if (t == 0) {
        s_init[0] = (int)atomicAdd(ctr, 1u) * nq + q;
        s_init[1] = (int)atomicAdd(ctr, 1u) * nq + q;
        s_init[2] = (int)atomicAdd(ctr, 1u) * nq + q;
        c_pend = (int)atomicAdd(ctr, 1u) * nq + q;
        st_pend = ld_relaxed(&sc->status);
    }
    __syncthreads();
    int c_cur = s_init[0], c_nxt = s_init[1], c_nxt2 = s_init[2];
    int4 m_cur = uniform(desc(c_cur));
    int4 m_nxt = uniform(desc(c_nxt));
    int j_cur[kBulkPref], j_nxt[kBulkPref];
    T l_cur[kBulkPref], l_nxt[kBulkPref];
    fetch(m_cur, j_cur, l_cur);
    // the first chunk's slices are waited for here, once: were they still in
    // flight at the loop entry, the compiler would wait at the top of every
    // iteration for loads issued that early, i.e. for the prefetch as well
    __builtin_amdgcn_s_waitcnt(0);
    int par = 0;
    while (c_cur < nb) {
        fetch(m_nxt, j_nxt, l_nxt);
        const int4 m_nxt2 = desc(c_nxt2);
        if (kFwd) fwd_chunk<T>(a, m_cur, t, alpha, p_cur, update_rx, j_cur, l_cur);
        else bwd_chunk<T>(a, m_cur, t, j_cur, l_cur);
        // made uniform before thread 0 issues new loads, so that waiting for
        // it does not also wait for them
        const int4 m_nxt2_u = uniform(m_nxt2);
        if (t == 0) {
            // publish the claim taken a chunk ago, take the next one; both
            // this atomic and the status load complete during the next chunk
            s_claim[par] = c_pend;              // slots alternate, so one barrier suffices
            s_stop = (int)st_pend;
            c_pend = (int)atomicAdd(ctr, 1u) * nq + q;
            st_pend = ld_relaxed(&sc->status);
        }
        __syncthreads();
        if (s_stop & kPcgTimeout) break;
        c_cur = c_nxt;
        c_nxt = c_nxt2;
        c_nxt2 = s_claim[par];
        m_cur = m_nxt;
        m_nxt = m_nxt2_u;
#pragma unroll
        for (int u = 0; u < kBulkPref; ++u) {
            j_cur[u] = j_nxt[u];
            l_cur[u] = l_nxt[u];
        }
        par ^= 1;
    }
}

// ---------------------------------------------------------------- tail
// The narrow trailing levels (rows [tail_pos0, n)) run in one kTailThreads
// workgroup whose LDS holds the tail's solution values, so the level-to-level
// dependency chain goes through LDS and a barrier, not through HBM.
// Forward: k_tail_prep folds every tail row's entries in the wide levels
// (final once the bulk sweep has ended) into tacc with one wave per row, and
// applies the row's r/x update; k_tail_solve then walks the tail levels.
// Backward: the tail levels come first and depend only on each other.
template <typename T>
__global__ void __launch_bounds__(kRedBlock) k_tail_prep(SolveArgs a, const double *p_cur, int update_rx)
{
    PcgScalars *sc = a.sc;
    if (sc->done) return;
    const double alpha0 = update_rx ? sc->alpha : 0.0;
    const T *val = (const T *)a.tp_val;
    const T *y = (const T *)a.y;
    T *tacc = (T *)a.tacc;
    const int lane = lane_id();
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * blockDim.x / kWave;
    for (int q = wave; q < a.ntail; q += nwaves) {
        T acc = (T)0;
        for (int e = a.tp_ptr[q] + lane; e < a.tp_ptr[q + 1]; e += kWave) acc = fma_t(-val[e], y[a.tp_ind[e]], acc);
        acc = group_sum(acc, kWave);
        if (lane == kWave - 1) {
            const int k = a.tail_pos0 + q;
            double rk = a.r[k];
            if (update_rx) {
                const double al = a.alpha_sys ? a.alpha_sys[a.sys_of[k]] : alpha0;
                rk = fma_t(-al, a.ap[k], rk);
                a.r[k] = rk;
                a.x[k] = fma_t(al, p_cur[k], a.x[k]);
            }
            tacc[q] = (T)rk + acc;
        }
    }
}

// Waits for this wave's LDS operations, then the workgroup barrier; global
// loads in flight (the next level's indices) are not waited for.
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

constexpr int kTailPref = kTailBlock;    // slices per lane loaded before the LDS reads

// The kTailBlock consecutive entries of a lane slot (16-byte aligned: every
// level's block size is a multiple of kTailBlock entries).
template <typename T>
__device__ __forceinline__ void tail_block_load(const int *ind, const T *val, int base, int *j, T *l)
{
    static_assert(kTailBlock == 8, "two 16-byte index loads per block");
    const int4 *pi = reinterpret_cast<const int4 *>(ind + base);
    const int4 i0 = pi[0], i1 = pi[1];
    j[0] = i0.x; j[1] = i0.y; j[2] = i0.z; j[3] = i0.w;
    j[4] = i1.x; j[5] = i1.y; j[6] = i1.z; j[7] = i1.w;
    if constexpr (sizeof(T) == 4) {
        const float4 *pv = reinterpret_cast<const float4 *>(val + base);
        const float4 v0 = pv[0], v1 = pv[1];
        l[0] = v0.x; l[1] = v0.y; l[2] = v0.z; l[3] = v0.w;
        l[4] = v1.x; l[5] = v1.y; l[6] = v1.z; l[7] = v1.w;
    } else {
        const double2 *pv = reinterpret_cast<const double2 *>(val + base);
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const double2 v = pv[h];
            l[2 * h] = v.x;
            l[2 * h + 1] = v.y;
        }
    }
}

template <typename T, bool kFwd>
__global__ void __launch_bounds__(kTailThreads) k_tail_solve(SolveArgs a)
{
    extern __shared__ double tail_lds_d[];
    T *v = reinterpret_cast<T *>(tail_lds_d);
    PcgScalars *sc = a.sc;
    if (sc->done) return;
    const int t = threadIdx.x;
    const int nt = a.ntail, tp0 = a.tail_pos0;
    T *y = (T *)a.y;
    T *z = (T *)a.z;
    const int *off = kFwd ? a.tf_off : a.tb_off;
    const int *slices = kFwd ? a.tf_slices : a.tb_slices;
    const int *ind = kFwd ? a.tf_ind : a.tb_ind;
    const T *val = (const T *)(kFwd ? a.tf_val : a.tb_val);
    // initial values: forward = bulk-folded right-hand side; backward = D+ y
    // (solver.hpp:64-84), with y reset to the sentinel for the next sweep
    for (int q = t; q < nt; q += kTailThreads) {
        T init;
        if (kFwd) {
            init = ((const T *)a.tacc)[q];
        } else {
            const int k = tp0 + q;
            T yk = y[k];
            y[k] = sent_value(T());
            const T dk = ((const T *)a.diag)[k];
            if (dk != (T)0) yk = yk / dk;
            init = yk;
        }
        v[q] = init;
    }
    if (t == 0) v[nt] = (T)0;        // slot of the padding entries
    // level descriptors {start, cnt, ELL offset, slices} after the values
    const int ntl = a.ntail_levels;
    const int ze = off[ntl];          // the padding block after the last level
    int4 *tm = reinterpret_cast<int4 *>(reinterpret_cast<char *>(tail_lds_d) + tail_value_bytes<T>(nt));
    for (int L = t; L < ntl; L += kTailThreads) tm[L] = make_int4(a.tl_start[L], a.tl_cnt[L], off[L], slices[L]);
    // an empty level (no rows, no slices) after the last one: the loop below
    // solves whole groups of three levels, the ones past the end being empty
    if (t == 0) tm[ntl] = make_int4(tp0, 0, ze, 0);
    __syncthreads();
    // Software pipeline over three register sets (A, B, C): this lane's first
    // kTailPref index/value slices of level s + 2 are loaded while level s is
    // solved (two levels of latency cover one global round trip); the loop
    // is unrolled by three so that no register copy (and no wait) sits
    // between the loads and their use.  Level descriptors are read from LDS
    // one fetch ahead (dn), so a fetch starts without an LDS round trip; the
    // lane term (a.zero is 0) keeps the read a per-lane one, so the compiler
    // does not make it uniform (and wait for it) at the read but at its use.
    auto desc = [&](int s) { return tm[(s < ntl ? (kFwd ? s : ntl - 1 - s) : ntl) + (t & (kWave - 1)) * a.zero]; };
    auto fetch = [&](const int4 &dsrc, int4 &d, int *j, T *l) {
        d = make_int4(__builtin_amdgcn_readfirstlane(dsrc.x), __builtin_amdgcn_readfirstlane(dsrc.y),
                      __builtin_amdgcn_readfirstlane(dsrc.z), __builtin_amdgcn_readfirstlane(dsrc.w));
        const int stride = sweep_rows_pow2(d.y) * tail_group(d.y);
        // every lane loads (the padding block when it has no slot), so the
        // loads are never predicated away, and nothing is computed from the
        // loaded values before their level is solved (a mask applied here
        // would make the compiler wait for these loads at once); slices past
        // the level's own are padding in the layout.  A level without
        // internal entries (d.w = 0, also the empty one past the end) has no
        // block of its own.
        tail_block_load<T>(ind, val, (t < stride && d.w > 0) ? d.z + t * kTailBlock : ze, j, l);
    };
    auto solve = [&](const int4 &d, const int *j, const T *l) {
        const int G = tail_group(d.y);
        const int stride = sweep_rows_pow2(d.y) * G;
        const int i = t / G, g = t - i * G;
        const bool lead = t < stride && i < d.y && g == G - 1;
        const int q = d.x - tp0 + i;
        T acc = (T)0;
        // a wave with no slot of this level skips its LDS work (its loads of
        // the padding block were issued all the same, see fetch)
        if ((t & ~(kWave - 1)) < stride) {
            // the row's own value is read with the gathers (it is written only
            // at the end of the level), not after the sum: one LDS round trip
            // less on the level's critical path; other lanes read the zero slot
            const T own = v[lead ? q : nt];
#pragma unroll
            for (int u = 0; u < kTailPref; ++u) acc = fma_t(-l[u], v[j[u]], acc);
            if (t < stride) {
                // blocks past the prefetched one (long rows); a padded entry
                // adds -0 * 0 = -0, which leaves acc unchanged
                for (int sl = kTailPref; sl < d.w; sl += kTailBlock) {
                    int jj[kTailBlock];
                    T ll[kTailBlock];
                    tail_block_load<T>(ind, val, d.z + sl * stride + t * kTailBlock, jj, ll);
#pragma unroll
                    for (int u = 0; u < kTailBlock; ++u) acc = fma_t(-ll[u], v[jj[u]], acc);
                }
            }
            if (lead) acc += own;
            acc = group_sum(acc, G);
            if (lead) v[q] = acc;
        }
        lds_barrier();
    };
    int4 dA = make_int4(0, 0, 0, 0), dB = dA, dC = dA;
    int jA[kTailPref], jB[kTailPref], jC[kTailPref];
    T lA[kTailPref], lB[kTailPref], lC[kTailPref];
    int4 dn = dA;
    if (ntl > 0) {
        fetch(desc(0), dA, jA, lA);
        fetch(desc(1), dB, jB, lB);
        dn = desc(2);
    }
    // no exits inside the group: an exit path would carry the newest loads
    // to the loop header, where the compiler then waits for them every level
    for (int s = 0; s < ntl; s += 3) {
        fetch(dn, dC, jC, lC);
        dn = desc(s + 3);
        solve(dA, jA, lA);
        fetch(dn, dA, jA, lA);
        dn = desc(s + 4);
        solve(dB, jB, lB);
        fetch(dn, dB, jB, lB);
        dn = desc(s + 5);
        solve(dC, jC, lC);
    }
    for (int q = t; q < nt; q += kTailThreads) {
        const int k = tp0 + q;
        if (kFwd) {
            y[k] = unsent(v[q]);
            z[k] = sent_value(T());
        } else {
            z[k] = unsent(v[q]);
        }
    }
}

// ---------------------------------------------------------------- K4
template <typename T>
__global__ void __launch_bounds__(kRedBlock) k_dot(SolveArgs a, int init)
{
    __shared__ double lds[kRedBlock / kWave];
    __shared__ int flag;
    PcgScalars *sc = a.sc;
    if (sc->done) return;
    const T *z = (const T *)a.z;
    const bool need_rr = sc->stop_rule == 1;
    double acc = 0.0, acc_rr = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
        double ri = a.r[i];
        acc += ri * (double)z[i];
        if (need_rr) acc_rr += ri * ri;
    }
    double bs = block_sum(acc, lds);
    double bs_rr = need_rr ? block_sum(acc_rr, lds) : 0.0;
    if (threadIdx.x == 0) {
        st_relaxed_val(a.partials + blockIdx.x, bs);
        st_relaxed_val(a.partials + gridDim.x + blockIdx.x, bs_rr);
    }
    if (last_block(&sc->arrive_dot, &flag)) {
        double rz = reduce_partials(a.partials, gridDim.x, lds);
        double rr = need_rr ? reduce_partials(a.partials + gridDim.x, gridDim.x, lds) : 0.0;
        if (threadIdx.x == 0) {
            bool go;
            if (init) {
                sc->rz0 = rz;
                sc->rz = rz;
                sc->beta = 0.0;
                sc->iter = 0;
                go = need_rr ? (sc->bb > 0.0 && sqrt(rr) / sqrt(sc->bb) > sc->tol)
                             : (rz > sc->tol);                         // solver.hpp:493
            } else {
                sc->beta = rz / sc->rz;
                sc->rz = rz;
                sc->iter += 1;
                go = need_rr ? (sqrt(rr) / sqrt(sc->bb) > sc->tol)
                             : (sqrt(rz) / sqrt(sc->rz0) > sc->tol);  // solver.hpp:536
            }
            sc->rr = rr;
            if (!go || sc->iter >= sc->maxit || sc->status != 0) sc->done = 1;
            sc->arrive_dot = 0u;
        }
    }
}

// ||A x - b||^2 and ||b||^2 (solver.hpp:549-569).
__global__ void __launch_bounds__(kRedBlock) k_residual(SolveArgs a)
{
    __shared__ double lds[kRedBlock / kWave];
    __shared__ int flag;
    PcgScalars *sc = a.sc;
    const int lane = lane_id();
    double acc = 0.0, accb = 0.0;
    const TileRange tr = xcd_tile_range(a.ntiles);
    for (int it = tr.first; it < tr.end; it += tr.stride) {
        const int tile = a.a_tile_ord[it / kWave];
        const int p = tile * kWave + lane;
        const int w = a.a_width[tile];
        const int *ind = a.a_ind + a.a_off[tile] + lane;
        const double *val = a.a_val + a.a_off[tile] + lane;
        double s = 0.0;
        for (int t = 0; t < w; ++t) s = fma_t(val[t * kWave], a.x[ind[t * kWave]], s);
        if (p < a.n) {
            double d = s - a.b[p];
            acc += d * d;
            accb += a.b[p] * a.b[p];
        }
    }
    double bs = block_sum(acc, lds);
    double bb = block_sum(accb, lds);
    if (threadIdx.x == 0) {
        st_relaxed_val(a.partials + blockIdx.x, bs);
        st_relaxed_val(a.partials + gridDim.x + blockIdx.x, bb);
    }
    if (last_block(&sc->arrive_res, &flag)) {
        double r2 = reduce_partials(a.partials, gridDim.x, lds);
        double b2 = reduce_partials(a.partials + gridDim.x, gridDim.x, lds);
        if (threadIdx.x == 0) {
            sc->res2 = r2;
            sc->bnorm2 = b2;
            sc->arrive_res = 0u;
        }
    }
}

template <typename T>
__global__ void k_fill_sentinel(T *v, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] = sent_value(T());
}

template <typename T>
__global__ void k_pcg_prepare(SolveArgs a)
{
    // r = b, x = 0, p = 0, Ap = 0 (solver.hpp:268-277, 462-466)
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
        a.r[i] = a.b[i];
        a.x[i] = 0.0;
        a.p0[i] = 0.0;
        a.p1[i] = 0.0;
        a.ap[i] = 0.0;
    }
    if (blockIdx.x == 0) reset_claims(a.ctr_fwd, a.ctr_bwd, threadIdx.x);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.x[a.n] = 0.0; a.r[a.n] = 0.0; a.p0[a.n] = 0.0; a.p1[a.n] = 0.0; a.ap[a.n] = 0.0;
        ((T *)a.y)[a.n] = (T)0;     // dummy slot read by ELL padding
        ((T *)a.z)[a.n] = (T)0;
    }
}

__global__ void k_bnorm(SolveArgs a)
{
    __shared__ double lds[kRedBlock / kWave];
    __shared__ int flag;
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) acc += a.b[i] * a.b[i];
    double bs = block_sum(acc, lds);
    if (threadIdx.x == 0) st_relaxed_val(a.partials + blockIdx.x, bs);
    if (last_block(&a.sc->arrive_res, &flag)) {
        double b2 = reduce_partials(a.partials, gridDim.x, lds);
        if (threadIdx.x == 0) { a.sc->bb = b2; a.sc->arrive_res = 0u; }
    }
}

__global__ void k_gather(const double *src, const int *idx, double *dst, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[idx[i]];
}
__global__ void k_scatter(const double *src, const int *idx, double *dst, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[idx[i]] = src[i];
}
// Plain streaming copy (calibration of bandwidth and of the HBM counters).
__global__ void __launch_bounds__(256) k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, size_t n2)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
__global__ void k_reset_counters(unsigned int *a, unsigned int *b)
{
    reset_claims(a, b, threadIdx.x);
}

int sweep_blocks(const SolveArgs &a)
{
    int g = a.bulk_chunks < a.sweep_blocks ? a.bulk_chunks : a.sweep_blocks;
    return g > 0 ? g : 1;
}

template <typename T>
void launch_fwd(const SolveArgs &a, const double *p_cur, int update_rx, hipStream_t s)
{
    if (a.bulk_chunks > 0)
        hipLaunchKernelGGL((k_sweep_bulk<T, true>), dim3(sweep_blocks(a)), dim3(kSweepThreads), 0, s, a, p_cur,
                           update_rx);
    if (a.ntail > 0) {
        int g = (a.ntail + 3) / 4;
        if (g > 4096) g = 4096;
        hipLaunchKernelGGL(k_tail_prep<T>, dim3(g), dim3(kRedBlock), 0, s, a, p_cur, update_rx);
        hipLaunchKernelGGL((k_tail_solve<T, true>), dim3(1), dim3(kTailThreads), a.tail_lds_bytes, s, a);
    }
}

template <typename T>
void launch_bwd(const SolveArgs &a, hipStream_t s)
{
    if (a.ntail > 0)
        hipLaunchKernelGGL((k_tail_solve<T, false>), dim3(1), dim3(kTailThreads), a.tail_lds_bytes, s, a);
    if (a.bulk_chunks > 0)
        hipLaunchKernelGGL((k_sweep_bulk<T, false>), dim3(sweep_blocks(a)), dim3(kSweepThreads), 0, s, a,
                           (const double *)nullptr, 0);
}

// ---------------------------------------------------------------- batched systems
// Block-diagonal system of nsys independent Laplacians (no entry couples two
// systems), one PCG per system.  The SpMV and the sweeps run over the whole
// batch in one launch each; alpha and beta are looked up per row through
// sys_of, and the per-system dot products are segmented reductions over the
// system-major position list (see BatchArgs).  A converged system is frozen:
// its alpha and beta are zero, so x and r keep their values and p stays z.
enum { kOpPap = 0, kOpRz = 1, kOpBnorm = 2, kOpRes = 3 };

// p_new = z + beta_sys p_old per row (streaming; the system's beta is looked
// up once per row instead of once per matrix entry in the SpMV gather).
template <typename T>
__global__ void __launch_bounds__(kRedBlock) k_bpupdate(SolveArgs a, BatchArgs b, const double *p_old, double *p_new)
{
    PcgScalars *sc = a.sc;
    if (sc->done) return;
    const T *z = (const T *)a.z;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x)
        p_new[i] = b.beta[b.sys_of[i]] * p_old[i] + (double)z[i];
}

// Ap = A p over the block-diagonal batch (every neighbour of a row belongs
// to the row's system).
__global__ void __launch_bounds__(kRedBlock) k_bspmv(SolveArgs a, const double *p)
{
    PcgScalars *sc = a.sc;
    if (sc->done) return;
    if (blockIdx.x == 0) reset_claims(a.ctr_fwd, a.ctr_bwd, threadIdx.x);
    const int lane = lane_id();
    const TileRange tr = xcd_tile_range(a.ntiles);
    for (int it = tr.first; it < tr.end; it += tr.stride) {
        const int tile = a.a_tile_ord[it / kWave];
        const int q = tile * kWave + lane;
        const int w = a.a_width[tile];
        const int *ind = a.a_ind + a.a_off[tile] + lane;
        const double *val = a.a_val + a.a_off[tile] + lane;
        double acc = 0.0;
        int t = 0;
        for (; t + 4 <= w; t += 4) {
            int j[4];
            double v[4], pj[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) { j[u] = ind[(t + u) * kWave]; v[u] = val[(t + u) * kWave]; }
#pragma unroll
            for (int u = 0; u < 4; ++u) pj[u] = p[j[u]];
#pragma unroll
            for (int u = 0; u < 4; ++u) acc = fma_t(v[u], pj[u], acc);
        }
        for (; t < w; ++t) acc = fma_t(val[t * kWave], p[ind[t * kWave]], acc);
        if (q < a.n) a.ap[q] = acc;
    }
}

// ap = A x - b (for the final residuals).
__global__ void __launch_bounds__(kRedBlock) k_bresid(SolveArgs a)
{
    const int lane = lane_id();
    const TileRange tr = xcd_tile_range(a.ntiles);
    for (int it = tr.first; it < tr.end; it += tr.stride) {
        const int tile = a.a_tile_ord[it / kWave];
        const int p = tile * kWave + lane;
        const int w = a.a_width[tile];
        const int *ind = a.a_ind + a.a_off[tile] + lane;
        const double *val = a.a_val + a.a_off[tile] + lane;
        double acc = 0.0;
        for (int t = 0; t < w; ++t) acc = fma_t(val[t * kWave], a.x[ind[t * kWave]], acc);
        if (p < a.n) a.ap[p] = acc - a.b[p];
    }
}

// One block per reduction block of the table (grid = nblk).  Op selects the
// sums (pap: p'Ap; rz: r'z and r'r; bnorm: b'b; res: d'd and b'b with d in ap)
// and the per-system update made by the system's last block.
template <typename T, int Op>
__global__ void __launch_bounds__(kRedBlock) k_bseg(SolveArgs a, BatchArgs b, const double *pv, int init)
{
    __shared__ double lds[kRedBlock / kWave];
    __shared__ int flag;
    PcgScalars *sc = a.sc;
    if ((Op == kOpPap || Op == kOpRz) && sc->done) return;
    const int blk = blockIdx.x;
    const int s = b.blk_sys[blk];
    const bool need_rr = sc->stop_rule == 1;
    const bool skip = (Op == kOpPap || (Op == kOpRz && !init)) && b.done[s] != 0;
    double acc = 0.0, acc2 = 0.0;
    if (!skip) {
        const T *z = (const T *)a.z;
        for (int q = b.blk_lo[blk] + threadIdx.x; q < b.blk_hi[blk]; q += blockDim.x) {
            const int p = b.spos[q];
            if (Op == kOpPap) {
                acc += pv[p] * a.ap[p];
            } else if (Op == kOpRz) {
                const double ri = a.r[p];
                acc += ri * (double)z[p];
                if (need_rr) acc2 += ri * ri;
            } else if (Op == kOpBnorm) {
                const double bi = a.b[p];
                acc += bi * bi;
            } else {
                const double d = a.ap[p], bi = a.b[p];
                acc += d * d;
                acc2 += bi * bi;
            }
        }
    }
    const double s1 = block_sum(acc, lds);
    const double s2 = block_sum(acc2, lds);
    if (threadIdx.x == 0) {
        st_relaxed_val(b.part + blk, s1);
        st_relaxed_val(b.part + b.nblk + blk, s2);
    }
    __syncthreads();
    const int b0 = b.sys_blk0[s], b1 = b.sys_blk0[s + 1];
    if (threadIdx.x == 0) {
        fence_release();
        const unsigned int old = atomicAdd(b.arrive + s, 1u);
        flag = (old == (unsigned int)(b1 - b0 - 1)) ? 1 : 0;
        if (flag) fence_acquire();
    }
    __syncthreads();
    if (!flag) return;
    double t1 = 0.0, t2 = 0.0;
    for (int i = b0 + threadIdx.x; i < b1; i += blockDim.x) {
        t1 += ld_relaxed_val(b.part + i);
        t2 += ld_relaxed_val(b.part + b.nblk + i);
    }
    t1 = block_sum(t1, lds);
    t2 = block_sum(t2, lds);
    if (threadIdx.x != 0) return;
    b.arrive[s] = 0u;
    const double tol = sc->tol;
    if (Op == kOpPap) {
        if (b.done[s]) {
            b.alpha[s] = 0.0;
        } else if (!(t1 > 0.0) || !isfinite(t1)) {        // breakdown of this system
            b.brk[s] = 1;
            b.done[s] = 1;
            b.alpha[s] = 0.0;
            b.beta[s] = 0.0;
        } else {
            b.alpha[s] = b.rz[s] / t1;
        }
        return;
    }
    if (Op == kOpBnorm) { b.bb[s] = t1; return; }
    if (Op == kOpRes) { b.res2[s] = t1; b.bn2[s] = t2; return; }
    // kOpRz: per-system beta and stop test (solver.hpp:493, 536), then the
    // last system to finish decides whether the batch is done
    if (init) {
        b.rz0[s] = t1;
        b.rz[s] = t1;
        b.rr[s] = t2;
        b.beta[s] = 0.0;
        b.alpha[s] = 0.0;
        b.iter[s] = 0;
        const bool go = need_rr ? (b.bb[s] > 0.0 && sqrt(t2) / sqrt(b.bb[s]) > tol) : (t1 > tol);
        b.done[s] = go ? 0 : 1;
    } else if (!b.done[s]) {
        b.beta[s] = t1 / b.rz[s];
        b.rz[s] = t1;
        b.rr[s] = t2;
        b.iter[s] += 1;
        const bool go = need_rr ? (sqrt(t2) / sqrt(b.bb[s]) > tol) : (sqrt(t1) / sqrt(b.rz0[s]) > tol);
        if (!go || b.iter[s] >= sc->maxit) {
            b.done[s] = 1;
            b.beta[s] = 0.0;
        }
    }
    fence_release();
    const unsigned int old = atomicAdd(b.arrive_all, 1u);
    if (old != (unsigned int)(b.nsys - 1)) return;
    fence_acquire();
    *b.arrive_all = 0u;
    int all = 1;
    for (int q = 0; q < b.nsys; ++q) all &= (ld_relaxed(b.done + q) != 0) ? 1 : 0;
    if (init) sc->iter = 0;
    else sc->iter += 1;
    if (all || sc->iter >= sc->maxit || sc->status != 0) sc->done = 1;
}

int batch_grid(const BatchArgs &b) { return b.nblk > 0 ? b.nblk : 1; }

}  // namespace

template <typename T>
void launch_fill_sentinel(T *v, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_sentinel<T>, dim3(2048), dim3(256), 0, s, v, n);
}

template <typename T>
void launch_pcg_init(const SolveArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_sentinel<T>, dim3(2048), dim3(256), 0, s, (T *)a.y, a.n);
    hipLaunchKernelGGL(k_fill_sentinel<T>, dim3(2048), dim3(256), 0, s, (T *)a.z, a.n);
    hipLaunchKernelGGL(k_pcg_prepare<T>, dim3(2048), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_bnorm, dim3(a.reduce_blocks), dim3(kRedBlock), 0, s, a);
    launch_fwd<T>(a, (const double *)a.p0, 0, s);
    launch_bwd<T>(a, s);
    hipLaunchKernelGGL(k_dot<T>, dim3(a.reduce_blocks), dim3(kRedBlock), 0, s, a, 1);
}

// parity 0: p0 -> p1, parity 1: p1 -> p0.  The first loop iteration uses
// parity 0 with beta forced to zero, which makes p = z (solver.hpp:477).
template <typename T>
void launch_pcg_iteration(const SolveArgs &a, int parity, hipStream_t s)
{
    double *p_old = parity ? a.p1 : a.p0;
    double *p_new = parity ? a.p0 : a.p1;
    hipLaunchKernelGGL(k_pupdate<T>, dim3(2048), dim3(kRedBlock), 0, s, a, (const double *)p_old, p_new);
    hipLaunchKernelGGL(k_spmv<T>, dim3(a.reduce_blocks), dim3(kRedBlock), 0, s, a, (const double *)p_new);
    launch_fwd<T>(a, (const double *)p_new, 1, s);
    launch_bwd<T>(a, s);
    hipLaunchKernelGGL(k_dot<T>, dim3(a.reduce_blocks), dim3(kRedBlock), 0, s, a, 0);
}

template <typename T>
void launch_spmv_only(const SolveArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_spmv_plain, dim3(a.reduce_blocks), dim3(kRedBlock), 0, s, a, (const double *)a.p0, a.ap);
}

template <typename T>
void launch_trsv_fwd_only(const SolveArgs &a, hipStream_t s)
{
    launch_fwd<T>(a, (const double *)a.p0, 0, s);
}

template <typename T>
void launch_trsv_bwd_only(const SolveArgs &a, hipStream_t s)
{
    launch_bwd<T>(a, s);
}

int tail_lds_limit(int device)
{
    int lim = 0;
    if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess || lim <= 0)
        lim = 64 * 1024;
    if (lim > 160 * 1024) lim = 160 * 1024;
    (void)hipFuncSetAttribute((const void *)k_tail_solve<float, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void *)k_tail_solve<float, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void *)k_tail_solve<double, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);
    (void)hipFuncSetAttribute((const void *)k_tail_solve<double, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              lim);
    return lim;
}

int sweep_occupancy_blocks(int device)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sweep_bulk<double, true>, kSweepThreads, 0) != hipSuccess)
        per_cu = 4;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) cus = 256;
    if (per_cu < 1) per_cu = 1;
    return per_cu * cus;
}

void launch_residual(const SolveArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_residual, dim3(a.reduce_blocks), dim3(kRedBlock), 0, s, a);
}

void launch_gather(const double *src, const int *idx, double *dst, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_gather, dim3(2048), dim3(256), 0, s, src, idx, dst, n);
}
void launch_scatter(const double *src, const int *idx, double *dst, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_scatter, dim3(2048), dim3(256), 0, s, src, idx, dst, n);
}
void launch_copy(const double *src, double *dst, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy, dim3(8192), dim3(256), 0, s, (const double2 *)src, (double2 *)dst, n / 2);
}
void launch_reset_counters(unsigned int *a, unsigned int *b, hipStream_t s)
{
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, a, b);
}

size_t batch_state_bytes(int nsys, int nblk)
{
    // 8 double arrays + 3 int arrays + arrive (nsys) + arrive_all, then 2 nblk partials
    return (size_t)nsys * (8 * sizeof(double) + 4 * sizeof(int)) + 64 + (size_t)2 * nblk * sizeof(double) + 256;
}

void batch_carve(BatchArgs &b, void *buf)
{
    char *c = (char *)buf;
    const size_t S = (size_t)b.nsys;
    double **d[8] = {&b.rz, &b.rz0, &b.alpha, &b.beta, &b.bb, &b.rr, &b.res2, &b.bn2};
    for (int i = 0; i < 8; ++i) { *d[i] = (double *)c; c += S * sizeof(double); }
    b.part = (double *)c;
    c += (size_t)2 * b.nblk * sizeof(double);
    b.iter = (int *)c; c += S * sizeof(int);
    b.done = (int *)c; c += S * sizeof(int);
    b.brk = (int *)c; c += S * sizeof(int);
    b.arrive = (unsigned int *)c; c += S * sizeof(unsigned int);
    b.arrive_all = (unsigned int *)c;
}

template <typename T>
void launch_batch_init(const SolveArgs &a, const BatchArgs &b, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_sentinel<T>, dim3(2048), dim3(256), 0, s, (T *)a.y, a.n);
    hipLaunchKernelGGL(k_fill_sentinel<T>, dim3(2048), dim3(256), 0, s, (T *)a.z, a.n);
    hipLaunchKernelGGL(k_pcg_prepare<T>, dim3(2048), dim3(256), 0, s, a);
    hipLaunchKernelGGL((k_bseg<T, kOpBnorm>), dim3(batch_grid(b)), dim3(kRedBlock), 0, s, a, b,
                       (const double *)nullptr, 1);
    launch_fwd<T>(a, (const double *)a.p0, 0, s);
    launch_bwd<T>(a, s);
    hipLaunchKernelGGL((k_bseg<T, kOpRz>), dim3(batch_grid(b)), dim3(kRedBlock), 0, s, a, b, (const double *)nullptr,
                       1);
}

template <typename T>
void launch_batch_iteration(const SolveArgs &a, const BatchArgs &b, int parity, hipStream_t s)
{
    double *p_old = parity ? a.p1 : a.p0;
    double *p_new = parity ? a.p0 : a.p1;
    hipLaunchKernelGGL(k_bpupdate<T>, dim3(a.reduce_blocks), dim3(kRedBlock), 0, s, a, b, (const double *)p_old, p_new);
    hipLaunchKernelGGL(k_bspmv, dim3(a.reduce_blocks), dim3(kRedBlock), 0, s, a, (const double *)p_new);
    hipLaunchKernelGGL((k_bseg<T, kOpPap>), dim3(batch_grid(b)), dim3(kRedBlock), 0, s, a, b, (const double *)p_new,
                       0);
    launch_fwd<T>(a, (const double *)p_new, 1, s);
    launch_bwd<T>(a, s);
    hipLaunchKernelGGL((k_bseg<T, kOpRz>), dim3(batch_grid(b)), dim3(kRedBlock), 0, s, a, b, (const double *)nullptr,
                       0);
}

void launch_batch_residual(const SolveArgs &a, const BatchArgs &b, hipStream_t s)
{
    hipLaunchKernelGGL(k_bresid, dim3(a.reduce_blocks), dim3(kRedBlock), 0, s, a);
    hipLaunchKernelGGL((k_bseg<double, kOpRes>), dim3(batch_grid(b)), dim3(kRedBlock), 0, s, a, b,
                       (const double *)nullptr, 0);
}

template void launch_batch_init<float>(const SolveArgs &, const BatchArgs &, hipStream_t);
template void launch_batch_init<double>(const SolveArgs &, const BatchArgs &, hipStream_t);
template void launch_batch_iteration<float>(const SolveArgs &, const BatchArgs &, int, hipStream_t);
template void launch_batch_iteration<double>(const SolveArgs &, const BatchArgs &, int, hipStream_t);
template void launch_fill_sentinel<float>(float *, int, hipStream_t);
template void launch_fill_sentinel<double>(double *, int, hipStream_t);
template void launch_pcg_init<float>(const SolveArgs &, hipStream_t);
template void launch_pcg_init<double>(const SolveArgs &, hipStream_t);
template void launch_pcg_iteration<float>(const SolveArgs &, int, hipStream_t);
template void launch_pcg_iteration<double>(const SolveArgs &, int, hipStream_t);
template void launch_spmv_only<float>(const SolveArgs &, hipStream_t);
template void launch_spmv_only<double>(const SolveArgs &, hipStream_t);
template void launch_trsv_fwd_only<float>(const SolveArgs &, hipStream_t);
template void launch_trsv_fwd_only<double>(const SolveArgs &, hipStream_t);
template void launch_trsv_bwd_only<float>(const SolveArgs &, hipStream_t);
template void launch_trsv_bwd_only<double>(const SolveArgs &, hipStream_t);

}  // namespace rc
