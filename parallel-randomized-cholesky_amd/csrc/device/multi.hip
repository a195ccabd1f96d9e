// multi.hip -- PCG for k <= 64 right-hand sides sharing one factor.
//
// Same recurrence as pcg.hip (solver.hpp:457-569) run independently for every
// right-hand side q: its own alpha, beta, r'z, stop test and iteration count;
// a converged right-hand side is frozen (its x and r no longer change) while
// the others continue.  Vectors are row-major n x K (K = next power of two
// >= k) in the level-order position numbering, so the K values of a row are
// contiguous: a wave handles 64/K rows with K lanes each, the index and value
// of every matrix entry are one broadcast load, and the gather of row j's K
// values is one coalesced access.  Matrices are CSR in position numbering.
// The sweeps keep the sentinel protocol of the single-RHS sweeps (per value).
#include "common.cuh"
#include "kernels.h"

namespace rc {

namespace {

constexpr int kMB = 256;   // threads per block

struct MultiView {
    int n, K, k;
    int lane, q, sub, rows_per_wave;
};

__device__ __forceinline__ bool q_done(const MultiArgs &a, int q) { return q >= a.k || a.sc->done[q]; }

template <typename T>
__device__ __noinline__ T mwait_slow(T *p, long long spin_ticks, MultiScalars *sc)
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
        for (int t = 0; t < nap; ++t) __builtin_amdgcn_s_sleep(2);
        nap = nap < 32 ? nap * 2 : 32;
    }
}

// ---------------------------------------------------------------- SpMM
// p_new = z + beta_q p_old (streaming); then Ap = A p_new and per-q partials
// of p'Ap.
template <typename T>
__global__ void __launch_bounds__(kMB) k_mpupdate(MultiArgs a, const double *p_old, double *p_new)
{
    MultiScalars *sc = a.sc;
    if (sc->all_done) return;
    const int K = a.K;
    const T *z = (const T *)a.z;
    const size_t total = (size_t)a.n * K;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int q = (int)(i & (size_t)(K - 1));
        const double beta = q < a.k ? sc->beta[q] : 0.0;
        p_new[i] = fma_t(beta, p_old[i], (double)z[i]);
    }
}

template <typename T>
__global__ void __launch_bounds__(kMB) k_mspmv(MultiArgs a, const double *p)
{
    __shared__ double red[kMB];
    __shared__ int flag;
    MultiScalars *sc = a.sc;
    if (sc->all_done) return;
    if (blockIdx.x == 0) reset_claims(a.ctr_fwd, a.ctr_bwd, threadIdx.x);
    const int K = a.K;
    const int q = threadIdx.x % K;
    const int rows_per_block = kMB / K;
    double acc_dot = 0.0;
    // rows visited in the input numbering (a row's neighbours are near it
    // there, not in positions), the workgroups of each XCD (workgroup b runs
    // on XCD b % 8) walking one contiguous eighth of it
    int r0 = blockIdx.x * rows_per_block + threadIdx.x / K, r1 = a.n, rstep = gridDim.x * rows_per_block;
    if (gridDim.x % 8 == 0) {
        const int x = blockIdx.x % 8, per = (a.n + 7) / 8;
        const int lo = min(a.n, x * per);
        r1 = min(a.n, lo + per);
        r0 = lo + (blockIdx.x / 8) * rows_per_block + threadIdx.x / K;
        rstep = (gridDim.x / 8) * rows_per_block;
    }
    // two rows per pass, the first kE entries of both loaded together, so the
    // dependent loads of one row (row index, row bounds, entries, gathers)
    // overlap with the other's; each row still sums its entries in order
    constexpr int kE = 8;
    for (int ru = r0; ru < r1; ru += 2 * rstep) {
        const bool two = ru + rstep < r1;
        int row[2], eb[2], ee[2];
        row[0] = a.row_ord[ru];
        row[1] = two ? a.row_ord[ru + rstep] : row[0];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            eb[h] = a.a_ptr[row[h]];
            ee[h] = a.a_ptr[row[h] + 1];
        }
        if (!two) ee[1] = eb[1];
        int j[2][kE];
        double v[2][kE], pj[2][kE];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int u = 0; u < kE; ++u) {
                const bool has = eb[h] + u < ee[h];
                j[h][u] = has ? a.a_ind[eb[h] + u] : 0;
                v[h][u] = has ? a.a_val[eb[h] + u] : 0.0;
            }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int u = 0; u < kE; ++u) pj[h][u] = p[(size_t)j[h][u] * K + q];
        double acc[2] = {0.0, 0.0};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int u = 0; u < kE; ++u)
                if (eb[h] + u < ee[h]) acc[h] = fma_t(v[h][u], pj[h][u], acc[h]);
            for (int e = eb[h] + kE; e < ee[h]; ++e) acc[h] = fma_t(a.a_val[e], p[(size_t)a.a_ind[e] * K + q], acc[h]);
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (h == 1 && !two) break;
            const size_t o = (size_t)row[h] * K + q;
            a.ap[o] = acc[h];
            acc_dot += p[o] * acc[h];
        }
    }
    // per-q block partial: threads t, t+K, t+2K, ... share q
    red[threadIdx.x] = acc_dot;
    __syncthreads();
    if ((int)threadIdx.x < K) {
        double s = 0.0;
        for (int t = threadIdx.x; t < kMB; t += K) s += red[t];
        st_relaxed_val(a.partials + (size_t)blockIdx.x * K + threadIdx.x, s);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        fence_release();
        unsigned int old = atomicAdd(&sc->arrive_spmv, 1u);
        flag = old == gridDim.x - 1;
        if (flag) fence_acquire();
    }
    __syncthreads();
    if (flag && (int)threadIdx.x < K) {
        const int qq = threadIdx.x;
        double pap = 0.0;
        for (unsigned int b = 0; b < gridDim.x; ++b) pap += ld_relaxed_val(a.partials + (size_t)b * K + qq);
        if (qq < a.k && !sc->done[qq]) {
            if (!(pap > 0.0) || !isfinite(pap)) {
                sc->alpha[qq] = 0.0;
                sc->done[qq] = 1;
                sc->breakdown[qq] = 1;
            } else {
                sc->alpha[qq] = sc->rz[qq] / pap;
            }
        } else if (qq < a.k) {
            sc->alpha[qq] = 0.0;
        }
        if (qq == 0) sc->arrive_spmv = 0u;
    }
}

// ---------------------------------------------------------------- sweeps
// Persistent: a block claims a chunk (<= 64 rows of one level) and solves its
// rows R at a time: L = K / V lanes per row, each lane V right-hand sides (V =
// 4 in FP32, 2 in FP64), so a gather of row j's values is one 16-byte access
// per lane; R = min(64, 1024 / L), so FP32 chunks take a single pass.  Entry
// indices and values are loaded eight at a time, then the eight gathers.
template <typename T, int V>
__device__ __forceinline__ void vload(const T *p, T *v)
{
    if constexpr (V == 4 && sizeof(T) == 4) {
        const float4 q = *reinterpret_cast<const float4 *>(p);
        v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    } else if constexpr (V >= 2 && sizeof(T) == 8) {
#pragma unroll
        for (int h = 0; h < V; h += 2) {
            const double2 q = *reinterpret_cast<const double2 *>(p + h);
            v[h] = q.x; v[h + 1] = q.y;
        }
    } else if constexpr (V == 2) {
        const float2 q = *reinterpret_cast<const float2 *>(p);
        v[0] = q.x; v[1] = q.y;
    } else {
        v[0] = *p;
    }
}

template <typename T, bool kFwd, int V>
__global__ void __launch_bounds__(1024) k_msweep(MultiArgs a, const double *p_cur, int update_rx)
{
    __shared__ int s_chunk, s_stop;
    MultiScalars *sc = a.sc;
    if (sc->all_done) return;
    const int K = a.K;
    const int L = K / V;
    const int t = threadIdx.x;
    const int i0 = t / L, g = t - i0 * L;
    const int R = blockDim.x / L;
    const int q0 = g * V;
    double alpha[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int q = q0 + v;
        alpha[v] = (kFwd && update_rx && q < a.k && !sc->done[q]) ? sc->alpha[q] : 0.0;
    }
    T *y = (T *)a.y;
    T *z = (T *)a.z;
    T *vec = kFwd ? y : z;
    const T *val = (const T *)(kFwd ? a.l_val : a.lt_val);
    const int *ptr = kFwd ? a.l_ptr : a.lt_ptr;
    const int *ind = kFwd ? a.l_ind : a.lt_ind;
    const T *diag = (const T *)a.diag;
    // claim queue of this workgroup: its j-th claim is chunk j * nq + cq
    const int nq = (int)gridDim.x < kClaimQueues ? (int)gridDim.x : kClaimQueues;
    const int cq = (int)(blockIdx.x % (unsigned int)nq);
    unsigned int *ctr = (kFwd ? a.ctr_fwd : a.ctr_bwd) + cq * kClaimStride;
    while (true) {
        if (t == 0) {
            s_chunk = (int)atomicAdd(ctr, 1u) * nq + cq;
            s_stop = (ld_relaxed(&sc->status) & kPcgTimeout) ? 1 : 0;
        }
        __syncthreads();
        int c = s_chunk;
        const int stop = s_stop;
        __syncthreads();
        if (c >= a.nchunks || stop) break;
        if (!kFwd) c = a.nchunks - 1 - c;
        const int cnt = a.chunk_cnt[c];
        for (int i = i0; i < cnt; i += R) {
        const int k = a.chunk_start[c] + i;
        const size_t o = (size_t)k * K + q0;
        T acc[V];
        if (kFwd) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                double rk = a.r[o + v];
                if (alpha[v] != 0.0) {
                    rk = fma_t(-alpha[v], a.ap[o + v], rk);
                    a.r[o + v] = rk;
                    a.x[o + v] = fma_t(alpha[v], p_cur[o + v], a.x[o + v]);
                }
                acc[v] = (T)rk;
            }
        } else {
            T yv[V];
            vload<T, V>(y + o, yv);
            const T dk = diag[k];
#pragma unroll
            for (int v = 0; v < V; ++v) {
                y[o + v] = sent_value(T());
                acc[v] = dk != (T)0 ? yv[v] / dk : yv[v];
            }
        }
        const int e1 = ptr[k + 1];
        int e = ptr[k];
        constexpr int B = 8;
        for (; e < e1; e += B) {
            int j[B];
            T l[B];
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const bool in = e + u < e1;
                j[u] = in ? ind[e + u] : a.n;            // dummy slot n holds zeros
                l[u] = in ? val[e + u] : (T)0;
            }
            T w[B][V];
#pragma unroll
            for (int u = 0; u < B; ++u) vload<T, V>(vec + (size_t)j[u] * K + q0, w[u]);
#pragma unroll
            for (int u = 0; u < B; ++u) {
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    T x = w[u][v];
                    if (is_sent(x)) x = mwait_slow(vec + (size_t)j[u] * K + q0 + v, a.spin_ticks, sc);
                    acc[v] = fma_t(-l[u], x, acc[v]);
                }
            }
        }
#pragma unroll
        for (int v = 0; v < V; ++v) {
            if (kFwd) {
                st_relaxed_val(y + o + v, unsent(acc[v]));
                z[o + v] = sent_value(T());
            } else {
                st_relaxed_val(z + o + v, unsent(acc[v]));
            }
        }
        }
    }
}

// ---------------------------------------------------------------- dots
template <typename T>
__global__ void __launch_bounds__(kMB) k_mdot(MultiArgs a, int init)
{
    __shared__ double red[kMB];
    __shared__ double red2[kMB];
    __shared__ int flag;
    MultiScalars *sc = a.sc;
    if (sc->all_done) return;
    const int K = a.K;
    const int q = threadIdx.x % K;
    const int rows_per_block = kMB / K;
    const T *z = (const T *)a.z;
    const bool need_rr = sc->stop_rule == 1;
    double acc = 0.0, acc_rr = 0.0;
    for (int row = blockIdx.x * rows_per_block + threadIdx.x / K; row < a.n; row += gridDim.x * rows_per_block) {
        const size_t o = (size_t)row * K + q;
        const double ri = a.r[o];
        acc += ri * (double)z[o];
        if (need_rr) acc_rr += ri * ri;
    }
    red[threadIdx.x] = acc;
    red2[threadIdx.x] = acc_rr;
    __syncthreads();
    if ((int)threadIdx.x < K) {
        double s = 0.0, s2 = 0.0;
        for (int t = threadIdx.x; t < kMB; t += K) { s += red[t]; s2 += red2[t]; }
        st_relaxed_val(a.partials + (size_t)blockIdx.x * K + threadIdx.x, s);
        st_relaxed_val(a.partials + (size_t)(gridDim.x + blockIdx.x) * K + threadIdx.x, s2);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        fence_release();
        unsigned int old = atomicAdd(&sc->arrive_dot, 1u);
        flag = old == gridDim.x - 1;
        if (flag) fence_acquire();
    }
    __syncthreads();
    if (!flag) return;
    if ((int)threadIdx.x < K) {
        const int qq = threadIdx.x;
        double rz = 0.0, rr = 0.0;
        for (unsigned int b = 0; b < gridDim.x; ++b) {
            rz += ld_relaxed_val(a.partials + (size_t)b * K + qq);
            rr += ld_relaxed_val(a.partials + (size_t)(gridDim.x + b) * K + qq);
        }
        if (qq < a.k) {
            if (need_rr && !(!init && sc->done[qq])) sc->rr[qq] = rr;
            if (init) {
                sc->rz0[qq] = rz;
                sc->rz[qq] = rz;
                sc->beta[qq] = 0.0;
                sc->iter[qq] = 0;
                bool go = need_rr ? (sc->bb[qq] > 0.0 && sqrt(rr) / sqrt(sc->bb[qq]) > sc->tol) : (rz > sc->tol);
                sc->done[qq] = go ? 0 : 1;
            } else if (!sc->done[qq]) {
                sc->beta[qq] = rz / sc->rz[qq];
                sc->rz[qq] = rz;
                sc->iter[qq] += 1;
                bool go = need_rr ? (sqrt(rr) / sqrt(sc->bb[qq]) > sc->tol) : (sqrt(rz) / sqrt(sc->rz0[qq]) > sc->tol);
                if (!go || sc->iter[qq] >= sc->maxit) sc->done[qq] = 1;
            } else {
                sc->beta[qq] = 0.0;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int all = 1;
        for (int qq = 0; qq < a.k; ++qq) all &= sc->done[qq] ? 1 : 0;
        if (sc->status != 0) all = 1;
        sc->all_done = all;
        sc->arrive_dot = 0u;
    }
}

// ||A x_q - b_q||^2 and ||b_q||^2 per q (solver.hpp:549-569)
__global__ void __launch_bounds__(kMB) k_mresidual(MultiArgs a, int bnorm_only)
{
    __shared__ double red[kMB];
    __shared__ double red2[kMB];
    __shared__ int flag;
    MultiScalars *sc = a.sc;
    const int K = a.K;
    const int q = threadIdx.x % K;
    const int rows_per_block = kMB / K;
    double acc = 0.0, accb = 0.0;
    for (int row = blockIdx.x * rows_per_block + threadIdx.x / K; row < a.n; row += gridDim.x * rows_per_block) {
        const size_t o = (size_t)row * K + q;
        if (!bnorm_only) {
            double s = 0.0;
            for (int e = a.a_ptr[row]; e < a.a_ptr[row + 1]; ++e)
                s = fma_t(a.a_val[e], a.x[(size_t)a.a_ind[e] * K + q], s);
            double d = s - a.b[o];
            acc += d * d;
        }
        accb += a.b[o] * a.b[o];
    }
    red[threadIdx.x] = acc;
    red2[threadIdx.x] = accb;
    __syncthreads();
    if ((int)threadIdx.x < K) {
        double s = 0.0, s2 = 0.0;
        for (int t = threadIdx.x; t < kMB; t += K) { s += red[t]; s2 += red2[t]; }
        st_relaxed_val(a.partials + (size_t)blockIdx.x * K + threadIdx.x, s);
        st_relaxed_val(a.partials + (size_t)(gridDim.x + blockIdx.x) * K + threadIdx.x, s2);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        fence_release();
        unsigned int old = atomicAdd(&sc->arrive_res, 1u);
        flag = old == gridDim.x - 1;
        if (flag) fence_acquire();
    }
    __syncthreads();
    if (flag && (int)threadIdx.x < K) {
        const int qq = threadIdx.x;
        double r2 = 0.0, b2 = 0.0;
        for (unsigned int b = 0; b < gridDim.x; ++b) {
            r2 += ld_relaxed_val(a.partials + (size_t)b * K + qq);
            b2 += ld_relaxed_val(a.partials + (size_t)(gridDim.x + b) * K + qq);
        }
        if (!bnorm_only) sc->res2[qq] = r2;
        sc->bb[qq] = b2;
        if (qq == 0) sc->arrive_res = 0u;
    }
}

template <typename T>
__global__ void k_mprepare(MultiArgs a)
{
    const size_t total = (size_t)(a.n + 1) * a.K;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const bool dummy = i >= (size_t)a.n * a.K;
        a.r[i] = dummy ? 0.0 : a.b[i];
        a.x[i] = 0.0;
        a.p0[i] = 0.0;
        a.p1[i] = 0.0;
        a.ap[i] = 0.0;
        ((T *)a.y)[i] = dummy ? (T)0 : sent_value(T());
        ((T *)a.z)[i] = dummy ? (T)0 : sent_value(T());
    }
    if (blockIdx.x == 0) reset_claims(a.ctr_fwd, a.ctr_bwd, threadIdx.x);
}

// b (positions, K-strided) from the user's column-major B (n x k):
// b[p*K + q] = B[q*n_user + idx_user[p]]
__global__ void k_mgather(const double *B, const int *idx_user, int n, int k, int K, double *b)
{
    const size_t total = (size_t)n * K;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int p = (int)(i / K), q = (int)(i % K);
        b[i] = q < k ? B[(size_t)q * n + idx_user[p]] : 0.0;
    }
}
__global__ void k_mscatter(const double *x, const int *idx_user, int n, int k, int K, double *X)
{
    const size_t total = (size_t)n * K;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const int p = (int)(i / K), q = (int)(i % K);
        if (q < k) X[(size_t)q * n + idx_user[p]] = x[i];
    }
}

int msweep_blocks(const MultiArgs &a)
{
    int g = a.nchunks < a.sweep_blocks ? a.nchunks : a.sweep_blocks;
    return g > 0 ? g : 1;
}

// One sweep: 64 rows x (K / V) lanes per block, V = 4 right-hand sides per lane
// (fewer when K < 4).
template <typename T, bool kFwd>
void launch_msweep(const MultiArgs &a, const double *p_cur, int update_rx, hipStream_t s)
{
    const int vmax = sizeof(T) == 4 ? 4 : 2;
    const int V = a.K >= vmax ? vmax : a.K;
    const int L = a.K / V;
    const dim3 block((L * 64 <= 1024 ? 64 : 1024 / L) * L);
    if constexpr (sizeof(T) == 4) {
        if (V == 4) {
            hipLaunchKernelGGL((k_msweep<T, kFwd, 4>), dim3(msweep_blocks(a)), block, 0, s, a, p_cur, update_rx);
            return;
        }
    }
    if (V == 2) hipLaunchKernelGGL((k_msweep<T, kFwd, 2>), dim3(msweep_blocks(a)), block, 0, s, a, p_cur, update_rx);
    else hipLaunchKernelGGL((k_msweep<T, kFwd, 1>), dim3(msweep_blocks(a)), block, 0, s, a, p_cur, update_rx);
}

}  // namespace

template <typename T>
void launch_multi_init(const MultiArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_mprepare<T>, dim3(2048), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_mresidual, dim3(a.reduce_blocks), dim3(kMB), 0, s, a, 1);
    launch_msweep<T, true>(a, (const double *)a.p0, 0, s);
    launch_msweep<T, false>(a, (const double *)nullptr, 0, s);
    hipLaunchKernelGGL(k_mdot<T>, dim3(a.reduce_blocks), dim3(kMB), 0, s, a, 1);
}

template <typename T>
void launch_multi_iteration(const MultiArgs &a, int parity, hipStream_t s)
{
    double *p_old = parity ? a.p1 : a.p0;
    double *p_new = parity ? a.p0 : a.p1;
    hipLaunchKernelGGL(k_mpupdate<T>, dim3(4096), dim3(kMB), 0, s, a, (const double *)p_old, p_new);
    hipLaunchKernelGGL(k_mspmv<T>, dim3(a.reduce_blocks), dim3(kMB), 0, s, a, (const double *)p_new);
    launch_msweep<T, true>(a, (const double *)p_new, 1, s);
    launch_msweep<T, false>(a, (const double *)nullptr, 0, s);
    hipLaunchKernelGGL(k_mdot<T>, dim3(a.reduce_blocks), dim3(kMB), 0, s, a, 0);
}

template <typename T>
__global__ void k_msentinel(T *v, size_t total)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x)
        v[i] = sent_value(T());
}

template <typename T>
void launch_multi_spmm_only(const MultiArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_mspmv<T>, dim3(a.reduce_blocks), dim3(kMB), 0, s, a, (const double *)a.p0);
}

template <typename T>
void launch_multi_fwd_only(const MultiArgs &a, hipStream_t s)
{
    launch_msweep<T, true>(a, (const double *)a.p0, 0, s);
}

template <typename T>
void launch_multi_bwd_only(const MultiArgs &a, hipStream_t s)
{
    launch_msweep<T, false>(a, (const double *)nullptr, 0, s);
}

template <typename T>
void launch_multi_sentinels(const MultiArgs &a, bool y, bool z, hipStream_t s)
{
    const size_t total = (size_t)a.n * a.K;         // the dummy row n keeps its zeros
    if (y) hipLaunchKernelGGL(k_msentinel<T>, dim3(2048), dim3(256), 0, s, (T *)a.y, total);
    if (z) hipLaunchKernelGGL(k_msentinel<T>, dim3(2048), dim3(256), 0, s, (T *)a.z, total);
}

void launch_multi_residual(const MultiArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_mresidual, dim3(a.reduce_blocks), dim3(kMB), 0, s, a, 0);
}

void launch_multi_gather(const double *B, const int *idx_user, int n, int k, int K, double *b, hipStream_t s)
{
    hipLaunchKernelGGL(k_mgather, dim3(2048), dim3(256), 0, s, B, idx_user, n, k, K, b);
}
void launch_multi_scatter(const double *x, const int *idx_user, int n, int k, int K, double *X, hipStream_t s)
{
    hipLaunchKernelGGL(k_mscatter, dim3(2048), dim3(256), 0, s, x, idx_user, n, k, K, X);
}

template void launch_multi_init<float>(const MultiArgs &, hipStream_t);
template void launch_multi_init<double>(const MultiArgs &, hipStream_t);
template void launch_multi_iteration<float>(const MultiArgs &, int, hipStream_t);
template void launch_multi_iteration<double>(const MultiArgs &, int, hipStream_t);
template void launch_multi_spmm_only<float>(const MultiArgs &, hipStream_t);
template void launch_multi_spmm_only<double>(const MultiArgs &, hipStream_t);
template void launch_multi_fwd_only<float>(const MultiArgs &, hipStream_t);
template void launch_multi_fwd_only<double>(const MultiArgs &, hipStream_t);
template void launch_multi_bwd_only<float>(const MultiArgs &, hipStream_t);
template void launch_multi_bwd_only<double>(const MultiArgs &, hipStream_t);
template void launch_multi_sentinels<float>(const MultiArgs &, bool, bool, hipStream_t);
template void launch_multi_sentinels<double>(const MultiArgs &, bool, bool, hipStream_t);

}  // namespace rc
