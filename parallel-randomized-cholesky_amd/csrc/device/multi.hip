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
// p_new = z + beta_q p_old ; Ap = A p_new ; partial p'Ap per q.
template <typename T>
__global__ void __launch_bounds__(kMB) k_mspmv(MultiArgs a, const double *p_old, double *p_new)
{
    __shared__ double red[kMB];
    __shared__ int flag;
    MultiScalars *sc = a.sc;
    if (sc->all_done) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) { *a.ctr_fwd = 0u; *a.ctr_bwd = 0u; }
    const int K = a.K;
    const int q = threadIdx.x % K;
    const double beta = q < a.k ? sc->beta[q] : 0.0;
    const T *z = (const T *)a.z;
    const int rows_per_block = kMB / K;
    double acc_dot = 0.0;
    for (int row = blockIdx.x * rows_per_block + threadIdx.x / K; row < a.n; row += gridDim.x * rows_per_block) {
        double acc = 0.0;
        for (int e = a.a_ptr[row]; e < a.a_ptr[row + 1]; ++e) {
            const int j = a.a_ind[e];
            const double pj = beta * p_old[(size_t)j * K + q] + (double)z[(size_t)j * K + q];
            acc = fma_t(a.a_val[e], pj, acc);
        }
        const size_t o = (size_t)row * K + q;
        const double pi = beta * p_old[o] + (double)z[o];
        p_new[o] = pi;
        a.ap[o] = acc;
        acc_dot += pi * acc;
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
// Persistent: a block claims a chunk of the level order and walks its rows,
// rows_per_block at a time (K lanes per row).
template <typename T, bool kFwd>
__global__ void __launch_bounds__(kMB) k_msweep(MultiArgs a, const double *p_cur, int update_rx)
{
    __shared__ int s_chunk, s_stop;
    MultiScalars *sc = a.sc;
    if (sc->all_done) return;
    const int K = a.K;
    const int q = threadIdx.x % K;
    const int sub = threadIdx.x / K;
    const int rows_per_block = kMB / K;
    const bool live = q < a.k;
    const double alpha = (kFwd && update_rx && live && !sc->done[q]) ? sc->alpha[q] : 0.0;
    T *y = (T *)a.y;
    T *z = (T *)a.z;
    const T *val = (const T *)(kFwd ? a.l_val : a.lt_val);
    const int *ptr = kFwd ? a.l_ptr : a.lt_ptr;
    const int *ind = kFwd ? a.l_ind : a.lt_ind;
    const T *diag = (const T *)a.diag;
    unsigned int *ctr = kFwd ? a.ctr_fwd : a.ctr_bwd;
    while (true) {
        if (threadIdx.x == 0) {
            s_chunk = (int)atomicAdd(ctr, 1u);
            s_stop = (ld_relaxed(&sc->status) & kPcgTimeout) ? 1 : 0;
        }
        __syncthreads();
        int c = s_chunk;
        const int stop = s_stop;
        __syncthreads();
        if (c >= a.nchunks || stop) break;
        if (!kFwd) c = a.nchunks - 1 - c;
        const int cnt = a.chunk_cnt[c];
        const int p0 = a.chunk_start[c];
        for (int i = sub; i < cnt; i += rows_per_block) {
            const int k = p0 + i;
            const size_t o = (size_t)k * K + q;
            T acc;
            if (kFwd) {
                double rk = a.r[o];
                if (update_rx && alpha != 0.0) {
                    rk = fma_t(-alpha, a.ap[o], rk);
                    a.r[o] = rk;
                    a.x[o] = fma_t(alpha, p_cur[o], a.x[o]);
                }
                acc = (T)rk;
            } else {
                T yk = y[o];
                y[o] = sent_value(T());
                T dk = diag[k];
                if (dk != (T)0) yk = yk / dk;
                acc = yk;
            }
            T *vec = kFwd ? y : z;
            const int e1 = ptr[k + 1];
            int e = ptr[k];
            for (; e + 4 <= e1; e += 4) {
                int j[4];
                T l[4], v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) { j[u] = ind[e + u]; l[u] = val[e + u]; }
#pragma unroll
                for (int u = 0; u < 4; ++u) v[u] = vec[(size_t)j[u] * K + q];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (is_sent(v[u])) v[u] = mwait_slow(vec + (size_t)j[u] * K + q, a.spin_ticks, sc);
                    acc = fma_t(-l[u], v[u], acc);
                }
            }
            for (; e < e1; ++e) {
                const int j = ind[e];
                T v = vec[(size_t)j * K + q];
                if (is_sent(v)) v = mwait_slow(vec + (size_t)j * K + q, a.spin_ticks, sc);
                acc = fma_t(-val[e], v, acc);
            }
            if (kFwd) {
                st_relaxed_val(y + o, unsent(acc));
                z[o] = sent_value(T());
            } else {
                st_relaxed_val(z + o, unsent(acc));
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
    if (blockIdx.x == 0 && threadIdx.x == 0) { *a.ctr_fwd = 0u; *a.ctr_bwd = 0u; }
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

}  // namespace

template <typename T>
void launch_multi_init(const MultiArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_mprepare<T>, dim3(2048), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_mresidual, dim3(a.reduce_blocks), dim3(kMB), 0, s, a, 1);
    hipLaunchKernelGGL((k_msweep<T, true>), dim3(msweep_blocks(a)), dim3(kMB), 0, s, a, (const double *)a.p0, 0);
    hipLaunchKernelGGL((k_msweep<T, false>), dim3(msweep_blocks(a)), dim3(kMB), 0, s, a, (const double *)nullptr, 0);
    hipLaunchKernelGGL(k_mdot<T>, dim3(a.reduce_blocks), dim3(kMB), 0, s, a, 1);
}

template <typename T>
void launch_multi_iteration(const MultiArgs &a, int parity, hipStream_t s)
{
    double *p_old = parity ? a.p1 : a.p0;
    double *p_new = parity ? a.p0 : a.p1;
    hipLaunchKernelGGL(k_mspmv<T>, dim3(a.reduce_blocks), dim3(kMB), 0, s, a, (const double *)p_old, p_new);
    hipLaunchKernelGGL((k_msweep<T, true>), dim3(msweep_blocks(a)), dim3(kMB), 0, s, a, (const double *)p_new, 1);
    hipLaunchKernelGGL((k_msweep<T, false>), dim3(msweep_blocks(a)), dim3(kMB), 0, s, a, (const double *)nullptr, 0);
    hipLaunchKernelGGL(k_mdot<T>, dim3(a.reduce_blocks), dim3(kMB), 0, s, a, 0);
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

}  // namespace rc
