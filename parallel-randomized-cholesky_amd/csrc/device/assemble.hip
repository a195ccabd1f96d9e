// assemble.hip -- factor compaction, transpose and level schedule.
//
// Replaces the reference's factor->CSR path (driver.cu:1310-1548, 1743-1870:
// construct_rowptr, remove_last_csr, shift_element, construct_csr,
// copy_diagonal, segment_sort_indices, CUB reduce/scan/segmented sort) and the
// SDDM ground trim (driver_physics.cu:1945-2010) with own kernels: a
// three-phase device scan, one compaction pass that also drops the ground row
// and converts to the solve precision, a counting-sort transpose followed by
// a per-row sort (so every row of L is in ascending column order and the
// forward-solve sums are deterministic), and a counting sort of rows by level
// for the triangular solves.
#include "common.cuh"
#include "kernels.h"

namespace rc {

namespace {
constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ int block_exclusive_scan(int v, int *lds_wave, int &total)
{
    // inclusive wave scan
    const int lane = lane_id();
    const int wid = threadIdx.x / kWave;
    int x = v;
    for (int off = 1; off < kWave; off <<= 1) {
        int y = __shfl_up(x, off, kWave);
        if (lane >= off) x += y;
    }
    if (lane == kWave - 1) lds_wave[wid] = x;
    __syncthreads();
    int carry = 0, tot = 0;
    for (int w = 0; w < kScanBlock / kWave; ++w) {
        if (w < wid) carry += lds_wave[w];
        tot += lds_wave[w];
    }
    __syncthreads();
    total = tot;
    return carry + x - v;
}

__global__ void __launch_bounds__(kScanBlock) k_scan_tiles(const int *in, int *out, int n, int *tile_sums)
{
    __shared__ int lds_wave[kScanBlock / kWave];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
    for (int t = 0; t < kScanItems; ++t) {
        int i = base + t;
        v[t] = (i < n) ? in[i] : 0;
        s += v[t];
    }
    int total;
    int ex = block_exclusive_scan(s, lds_wave, total);
    for (int t = 0; t < kScanItems; ++t) {
        int i = base + t;
        if (i < n) out[i] = ex;
        ex += v[t];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// One workgroup scans the tile sums in place (exclusive) and writes the total.
__global__ void __launch_bounds__(kScanBlock) k_scan_single(int *sums, int m, int *total_out)
{
    __shared__ int lds_wave[kScanBlock / kWave];
    int carry = 0;
    for (int base = 0; base < m; base += kScanBlock) {
        int i = base + threadIdx.x;
        int v = (i < m) ? sums[i] : 0;
        int total;
        int ex = block_exclusive_scan(v, lds_wave, total);
        if (i < m) sums[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) *total_out = carry;
}

__global__ void k_scan_add(int *out, int n, const int *tile_off)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] += tile_off[i / kScanTile];
}
}  // namespace

size_t scan_tmp_elems(int n) { return (size_t)(n + kScanTile - 1) / kScanTile + 1; }

void scan_exclusive(const int *in, int *out, int n, int *tmp, hipStream_t s)
{
    int tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles < 1) tiles = 1;
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kScanBlock), 0, s, in, out, n, tmp);
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kScanBlock), 0, s, tmp, tiles, out + n);
    hipLaunchKernelGGL(k_scan_add, dim3(1024), dim3(256), 0, s, out, n, (const int *)tmp);
}

// ---------------------------------------------------------------- Lt
namespace {
__global__ void k_lt_counts(AssembleArgs a, int *cnt)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < a.ns; k += gridDim.x * blockDim.x) {
        int c = a.out_cnt[k];
        if (a.trim && c > 0 && a.out_row[a.out_start[k] + c - 1] == a.n - 1) --c;
        cnt[k] = c;
    }
}

template <typename T>
__global__ void k_lt_fill(AssembleArgs a, const int *lt_ptr, int *lt_ind, T *lt_val, T *diag_out)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < a.ns; k += gridDim.x * blockDim.x) {
        long long src = a.out_start[k];
        int dst = lt_ptr[k], len = lt_ptr[k + 1] - dst;
        for (int t = 0; t < len; ++t) {
            lt_ind[dst + t] = a.out_row[src + t];
            lt_val[dst + t] = (T)a.out_val[src + t];
        }
        diag_out[k] = (T)a.diag_in[k];
    }
}

__global__ void k_count_rows(int nnz, const int *ind, int *cnt)
{
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += gridDim.x * blockDim.x)
        atomicAdd(&cnt[ind[e]], 1);
}

template <typename T>
__global__ void k_transpose_scatter(int ns, const int *lt_ptr, const int *lt_ind, const T *lt_val, int *pos,
                                    int *l_ind, T *l_val)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < ns; k += gridDim.x * blockDim.x) {
        for (int e = lt_ptr[k]; e < lt_ptr[k + 1]; ++e) {
            int r = lt_ind[e];
            int p = atomicAdd(&pos[r], 1);
            l_ind[p] = k;
            l_val[p] = lt_val[e];
        }
    }
}

// One wave per segment.  Keys within a segment are distinct column indices.
// Segments up to kSegLds entries are sorted in LDS, longer ones in place in
// global memory; both with the ascending bitonic network (no padding).
constexpr int kSegLds = 1024;

template <typename T>
__device__ void seg_network(int *key, T *val, int len)
{
    const int lane = lane_id();
    int P = 1;
    while (P < len) P <<= 1;
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = lane; t < (P >> 1); t += kWave) {
                int i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
                int l = (j == (k >> 1)) ? (i ^ (k - 1)) : (i + j);
                if (l < len && key[i] > key[l]) {
                    int ki = key[i]; key[i] = key[l]; key[l] = ki;
                    T vi = val[i]; val[i] = val[l]; val[l] = vi;
                }
            }
            __syncthreads();
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kWave) k_segmented_sort(int nseg, const int *ptr, int *ind, T *val)
{
    __shared__ int skey[kSegLds];
    __shared__ T sval[kSegLds];
    const int lane = lane_id();
    for (int sgm = blockIdx.x; sgm < nseg; sgm += gridDim.x) {
        int b = ptr[sgm], len = ptr[sgm + 1] - b;
        if (len <= 1) continue;
        bool sorted = true;
        for (int t = lane; t + 1 < len; t += kWave)
            if (ind[b + t] > ind[b + t + 1]) sorted = false;
        if (!__any(!sorted)) continue;
        if (len <= kSegLds) {
            for (int t = lane; t < len; t += kWave) { skey[t] = ind[b + t]; sval[t] = val[b + t]; }
            __syncthreads();
            seg_network<T>(skey, sval, len);
            for (int t = lane; t < len; t += kWave) { ind[b + t] = skey[t]; val[b + t] = sval[t]; }
            __syncthreads();
        } else {
            seg_network<T>(ind + b, val + b, len);
        }
    }
}

__global__ void k_level_max(int ns, const int *level, unsigned long long *out)
{
    int m = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += gridDim.x * blockDim.x)
        m = level[i] > m ? level[i] : m;
    atomicMax(out, (unsigned long long)m);
}

__global__ void k_level_count(int ns, const int *level, int *cnt)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += gridDim.x * blockDim.x)
        atomicAdd(&cnt[level[i]], 1);
}

__global__ void k_level_scatter(int ns, const int *level, int *pos, int *order)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += gridDim.x * blockDim.x) {
        int p = atomicAdd(&pos[level[i]], 1);
        order[p] = i;
    }
}

__global__ void k_chunk_count(int nlevels, const int *lvl_ptr, int *cnt)
{
    for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < nlevels; l += gridDim.x * blockDim.x)
        cnt[l] = (lvl_ptr[l + 1] - lvl_ptr[l] + kWave - 1) / kWave;
}

__global__ void k_chunk_fill(int nchunks, int nlevels, const int *lvl_ptr, const int *chunk_ptr, int *chunk_start,
                             int *chunk_cnt)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nchunks; c += gridDim.x * blockDim.x) {
        int lo = 0, hi = nlevels - 1;          // last level with chunk_ptr[l] <= c
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (chunk_ptr[mid] <= c) lo = mid; else hi = mid - 1;
        }
        int start = lvl_ptr[lo] + (c - chunk_ptr[lo]) * kWave;
        int end = lvl_ptr[lo + 1];
        chunk_start[c] = start;
        chunk_cnt[c] = (end - start) < kWave ? (end - start) : kWave;
    }
}
}  // namespace

void launch_lt_counts(const AssembleArgs &a, int *cnt, hipStream_t s)
{
    hipLaunchKernelGGL(k_lt_counts, dim3(1024), dim3(256), 0, s, a, cnt);
}

template <typename T>
void launch_lt_fill(const AssembleArgs &a, const int *lt_ptr, int *lt_ind, T *lt_val, T *diag_out, hipStream_t s)
{
    hipLaunchKernelGGL(k_lt_fill<T>, dim3(2048), dim3(256), 0, s, a, lt_ptr, lt_ind, lt_val, diag_out);
}

void launch_count_rows(int nnz, const int *ind, int *cnt, hipStream_t s)
{
    hipLaunchKernelGGL(k_count_rows, dim3(2048), dim3(256), 0, s, nnz, ind, cnt);
}

template <typename T>
void launch_transpose_scatter(int ns, const int *lt_ptr, const int *lt_ind, const T *lt_val, int *pos, int *l_ind,
                              T *l_val, hipStream_t s)
{
    hipLaunchKernelGGL(k_transpose_scatter<T>, dim3(2048), dim3(256), 0, s, ns, lt_ptr, lt_ind, lt_val, pos, l_ind,
                       l_val);
}

template <typename T>
void launch_segmented_sort(int nseg, const int *ptr, int *ind, T *val, hipStream_t s)
{
    int g = nseg < 8192 ? (nseg > 0 ? nseg : 1) : 8192;
    hipLaunchKernelGGL(k_segmented_sort<T>, dim3(g), dim3(kWave), 0, s, nseg, ptr, ind, val);
}

void launch_level_max(int ns, const int *level, unsigned long long *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_level_max, dim3(1024), dim3(256), 0, s, ns, level, out);
}
void launch_level_count(int ns, const int *level, int *cnt, hipStream_t s)
{
    hipLaunchKernelGGL(k_level_count, dim3(1024), dim3(256), 0, s, ns, level, cnt);
}
void launch_level_scatter(int ns, const int *level, int *pos, int *order, hipStream_t s)
{
    hipLaunchKernelGGL(k_level_scatter, dim3(1024), dim3(256), 0, s, ns, level, pos, order);
}
void launch_chunk_count(int nlevels, const int *lvl_ptr, int *cnt, hipStream_t s)
{
    hipLaunchKernelGGL(k_chunk_count, dim3(256), dim3(256), 0, s, nlevels, lvl_ptr, cnt);
}
void launch_chunk_fill(int nchunks, int nlevels, const int *lvl_ptr, const int *chunk_ptr, int *chunk_start,
                       int *chunk_cnt, hipStream_t s)
{
    hipLaunchKernelGGL(k_chunk_fill, dim3(1024), dim3(256), 0, s, nchunks, nlevels, lvl_ptr, chunk_ptr, chunk_start,
                       chunk_cnt);
}

template void launch_lt_fill<float>(const AssembleArgs &, const int *, int *, float *, float *, hipStream_t);
template void launch_lt_fill<double>(const AssembleArgs &, const int *, int *, double *, double *, hipStream_t);
template void launch_transpose_scatter<float>(int, const int *, const int *, const float *, int *, int *, float *,
                                              hipStream_t);
template void launch_transpose_scatter<double>(int, const int *, const int *, const double *, int *, int *, double *,
                                               hipStream_t);
template void launch_segmented_sort<float>(int, const int *, int *, float *, hipStream_t);
template void launch_segmented_sort<double>(int, const int *, int *, double *, hipStream_t);

}  // namespace rc
