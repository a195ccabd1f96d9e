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
// rows in increasing order, one row per lane.  A row reads the rows it
// depends on (lower indices) once their pivot is published (pivots start as a
// sentinel and are stored last, after the row's values and level).  Lanes
// never block: every pass of a wave's loop, each lane whose next dependency
// is published consumes it, the others poll again on the next pass, so a
// lane waiting for a row of its own wave never holds that row back (a chunk
// processed by one lane row after row would chain every chunk to the last
// rows of the chunks before it).  A claimed row's dependencies were claimed
// earlier by running waves, so the smallest unfinished row can always
// proceed.  A wave that makes no progress sleeps briefly between passes; its
// waiting is bounded.
#include "common.cuh"
#include "kernels.h"

namespace rc {

namespace {

constexpr int kIcChunk = 64;

__global__ void __launch_bounds__(64) k_ic0(Ic0Args a)
{
    const int lane = threadIdx.x;
    while (true) {
        int c0 = 0;
        if (lane == 0) c0 = (int)atomicAdd(a.claim, 1u) * kIcChunk;
        c0 = __shfl(c0, 0, 64);
        if (c0 >= a.n) return;
        const int i = c0 + lane;
        bool done = i >= a.n;
        int b = 0, e = 0, t = 0, lev = 0;
        double di = 0.0;
        if (!done && i >= a.n_compute) {
            // the appended ground row of an SDDM input: trimmed from the
            // factor, and no row depends on the last one
            a.level[i] = 0;
            st_relaxed_val(a.d + i, 0.0);
            done = true;
        }
        if (!done) {
            b = a.up_ptr[i];
            e = a.up_ptr[i + 1];
            t = b;
            di = a.diag_a[i];
        }
        const long long t0 = wall_clock64();
        while (true) {
            bool progressed = false;
            if (!done && t < e) {
                const int j = a.up_ind[t];
                const double dj = ld_relaxed_val(a.d + j);
                if (!is_sent(dj)) {
                    fence_acquire();
                    // s = a_ij - sum over common m < j of l_im l_jm d_m, m
                    // ascending: every m of row j (short) is looked up in
                    // row i's entries before t by bisection, so a dense row
                    // costs O(log) per common neighbour, not a walk of its
                    // whole prefix
                    double sum = a.up_val[t];
                    int lo = b;
                    for (int q = a.up_ptr[j]; q < a.up_ptr[j + 1]; ++q) {
                        const int mq = a.up_ind[q];
                        int hi = t;                       // first p in [lo, t) with up_ind[p] >= mq
                        while (lo < hi) {
                            const int mid = (lo + hi) >> 1;
                            if (a.up_ind[mid] < mq) lo = mid + 1; else hi = mid;
                        }
                        if (lo < t && a.up_ind[lo] == mq) {
                            sum -= a.l[lo] * a.l[q] * a.d[mq];
                            ++lo;
                        }
                    }
                    const double lij = dj > 0.0 ? sum / dj : 0.0;
                    a.l[t] = lij;
                    di -= lij * lij * dj;
                    lev = max(lev, a.level[j] + 1);
                    ++t;
                    progressed = true;
                }
            }
            if (!done && t >= e) {
                if (!(di > 1e-14 * a.diag_a[i])) di = 0.0;
                a.level[i] = lev;
                fence_release();
                st_relaxed_val(a.d + i, di);
                done = true;
                progressed = true;
            }
            if (__all(done)) break;
            if (!__any(progressed)) {
                if ((ld_relaxed(a.status) & 1u) || wall_clock64() - t0 > a.spin_ticks) {
                    if (lane == 0) atomicOr(a.status, 1u);
                    return;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
    }
}

// L by columns in the elimination's output layout: row i's entry l_ij goes to
// the position of row i in column j of the strict-lower slice (rows sorted);
// out_val holds the entries of L, as the randomized elimination's -w/S do.
// Rows below n only (n excludes a trimmed ground row, whose entries the
// assembly drops).
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
