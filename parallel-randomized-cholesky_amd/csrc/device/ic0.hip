// ic0.hip -- incomplete Cholesky IC(0) of the prepared Laplacian, L D L^T with
// unit lower L on the pattern of the matrix's lower triangle: the context
// preconditioner the reference's experiments compare against (MATLAB ichol,
// experiment/ichol_graph_test.m:51-78; hypre / AmgX runs in hypre/ and amgx/).
// Written from the textbook recurrence, not from any library:
//   l_ij = (a_ij - sum_{m < j} l_im l_jm d_m) / d_j      (j < i, a_ij != 0)
//   d_i  = a_ii - sum_{j < i} l_ij^2 d_j
// with l_ij = 0 where d_j <= 0 and d_i clamped to 0 when rounding leaves it
// at or below 1e-14 a_ii (the zero pivot of a connected Laplacian), so the
// preconditioner's D+ rule skips it as it does the randomized factor's.
//
// Rows are computed by persistent waves that claim chunks of 64 consecutive
// rows in increasing order; a row reads the rows it depends on (lower
// indices) once their pivot is published (pivots start as a sentinel and are
// stored last, after the row's values and level).  A claimed row's
// dependencies were claimed earlier by running waves, so the smallest
// unfinished row can always proceed.  Every wait is bounded.
#include "common.cuh"
#include "kernels.h"

namespace rc {

namespace {

constexpr int kIcChunk = 64;

// Rows c0 .. c0+63 by one lane; returns 1 when a wait expired.
__device__ __noinline__ int ic0_chunk(const Ic0Args &a, int c0)
{
    const int c1 = min(a.n, c0 + kIcChunk);
    for (int i = c0; i < c1; ++i) {
        const int b = a.up_ptr[i], e = a.up_ptr[i + 1];
        double di = a.diag_a[i];
        int lev = 0;
        for (int t = b; t < e; ++t) {
            const int j = a.up_ind[t];
            // wait for row j (its pivot is published last)
            double dj = ld_relaxed_val(a.d + j);
            if (is_sent(dj)) {
                const long long t0 = wall_clock64();
                while (is_sent(dj = ld_relaxed_val(a.d + j))) {
                    if ((ld_relaxed(a.status) & 1u) || wall_clock64() - t0 > a.spin_ticks) {
                        atomicOr(a.status, 1u);
                        return 1;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
            fence_acquire();
            // s = a_ij - sum over common m < j of l_im l_jm d_m (both rows sorted)
            double s = a.up_val[t];
            int p = b, q = a.up_ptr[j];
            const int qe = a.up_ptr[j + 1];
            while (p < t && q < qe) {
                const int mp = a.up_ind[p], mq = a.up_ind[q];
                if (mp == mq) {
                    s -= a.l[p] * a.l[q] * a.d[mp];
                    ++p;
                    ++q;
                } else if (mp < mq) {
                    ++p;
                } else {
                    ++q;
                }
            }
            const double lij = dj > 0.0 ? s / dj : 0.0;
            a.l[t] = lij;
            di -= lij * lij * dj;
            lev = max(lev, a.level[j] + 1);
        }
        if (!(di > 1e-14 * a.diag_a[i])) di = 0.0;
        a.level[i] = lev;
        fence_release();
        st_relaxed_val(a.d + i, di);
    }
    return 0;
}

__global__ void __launch_bounds__(64) k_ic0(Ic0Args a)
{
    const int lane = threadIdx.x;
    while (true) {
        int c0 = 0;
        if (lane == 0) c0 = (int)atomicAdd(a.claim, 1u) * kIcChunk;
        c0 = __shfl(c0, 0, 64);
        if (c0 >= a.n) return;
        int stop = 0;
        if (lane == 0) stop = ic0_chunk(a, c0);
        stop = __shfl(stop, 0, 64);
        if (stop) return;
    }
}

// L by columns in the elimination's output layout: row i's entry l_ij goes to
// the position of row i in column j of the strict-lower slice (rows sorted);
// out_val holds the entries of L, as the randomized elimination's -w/S do.
__global__ void k_ic0_scatter(int n, const int *up_ptr, const int *up_ind, const double *l, const int *lo_ptr,
                              const int *lo_row, double *out_val)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        for (int t = up_ptr[i]; t < up_ptr[i + 1]; ++t) {
            const int j = up_ind[t];
            int lo = lo_ptr[j], hi = lo_ptr[j + 1];
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (lo_row[mid] < i) lo = mid + 1; else hi = mid;
            }
            out_val[lo] = l[t];
        }
    }
}

// Upper slice (rows j < i of column i) and the diagonal of every column of
// the prepared matrix, one wave per column (rows sorted: the upper entries
// are the column's first up_ptr[i+1] - up_ptr[i] entries).
__global__ void k_ic0_upper(int n, const int *ptr, const int *ind, const double *val, const int *up_ptr, int *up_ind,
                            double *up_val, double *diag_a)
{
    const int lane = threadIdx.x & 63;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    for (int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < n; i += nw) {
        const int b = ptr[i], len = up_ptr[i + 1] - up_ptr[i], o = up_ptr[i];
        for (int t = lane; t < len; t += 64) {
            up_ind[o + t] = ind[b + t];
            up_val[o + t] = val[b + t];
        }
        double dsum = 0.0;
        for (int p = b + lane; p < ptr[i + 1]; p += 64)
            if (ind[p] == i) dsum += val[p];
        dsum = wave_sum(dsum);
        if (lane == 0) diag_a[i] = dsum;
    }
}

__global__ void k_ptr_to_start_cnt(int n, const int *lo_ptr, long long *out_start, int *out_cnt)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        out_start[k] = lo_ptr[k];
        out_cnt[k] = lo_ptr[k + 1] - lo_ptr[k];
    }
}

}  // namespace

void launch_ptr_to_start_cnt(int n, const int *lo_ptr, long long *out_start, int *out_cnt, hipStream_t s)
{
    hipLaunchKernelGGL(k_ptr_to_start_cnt, dim3(2048), dim3(256), 0, s, n, lo_ptr, out_start, out_cnt);
}

void launch_ic0_upper(int n, const int *ptr, const int *ind, const double *val, const int *up_ptr, int *up_ind,
                      double *up_val, double *diag_a, hipStream_t s)
{
    hipLaunchKernelGGL(k_ic0_upper, dim3(4096), dim3(256), 0, s, n, ptr, ind, val, up_ptr, up_ind, up_val, diag_a);
}

void launch_ic0(const Ic0Args &a, int waves, hipStream_t s)
{
    hipLaunchKernelGGL(k_ic0, dim3(waves), dim3(64), 0, s, a);
}

void launch_ic0_scatter(int n, const int *up_ptr, const int *up_ind, const double *l, const int *lo_ptr,
                        const int *lo_row, double *out_val, hipStream_t s)
{
    hipLaunchKernelGGL(k_ic0_scatter, dim3(4096), dim3(256), 0, s, n, up_ptr, up_ind, l, lo_ptr, lo_row, out_val);
}

}  // namespace rc
