// assemble.hip -- factor compaction, transpose, level schedule, sliced ELL.
//
// Replaces the reference's factor->CSR path (driver.cu:1310-1548, 1743-1870:
// construct_rowptr, remove_last_csr, shift_element, construct_csr,
// copy_diagonal, segment_sort_indices, CUB reduce/scan/segmented sort) and the
// SDDM ground trim (driver_physics.cu:1945-2010) with own kernels:
//   * a three-phase device scan and one compaction pass (CSR of L^T, ground
//     row dropped, values converted to the solve precision);
//   * the transpose to rows of L by a stable radix sort of the entries by row
//     (each row ends up in ascending column order, so forward-solve sums are
//     deterministic and equal to the oracle's);
//   * a stable radix sort of rows by (level, descending row length): the
//     position order in which the triangular solves run and in which every
//     PCG vector is stored, so a 64-row chunk touches contiguous memory;
//   * sliced ELL (64 rows per slice, entry-major) for L, L^T and A so every
//     index/value load of a wave is one coalesced 256-byte access.
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
// One 16-lane group per row: the row's entries are read and written as
// contiguous runs.
__global__ void k_lt_fill(AssembleArgs a, const int *lt_ptr, int *lt_ind, T *lt_val, T *diag_out)
{
    const int g = threadIdx.x & 15;
    const int ngroups = (gridDim.x * blockDim.x) >> 4;
    for (int k = (blockIdx.x * blockDim.x + threadIdx.x) >> 4; k < a.ns; k += ngroups) {
        const long long src = a.out_start[k];
        const int dst = lt_ptr[k], len = lt_ptr[k + 1] - dst;
        for (int t = g; t < len; t += 16) {
            lt_ind[dst + t] = a.out_row[src + t];
            lt_val[dst + t] = (T)a.out_val[src + t];
        }
        if (g == 0) diag_out[k] = (T)a.diag_in[k];
    }
}

__global__ void k_level_max(int ns, const int *level, unsigned long long *out)
{
    int m = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += gridDim.x * blockDim.x)
        m = level[i] > m ? level[i] : m;
    for (int off = 32; off > 0; off >>= 1) {
        int o = __shfl_xor(m, off, kWave);
        m = o > m ? o : m;
    }
    if (lane_id() == 0) atomicMax(out, (unsigned long long)m);
}

__global__ void k_iota(int *v, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] = i;
}

__global__ void k_copy_keys(const int *src, uint32_t *dst, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = (uint32_t)src[i];
}

// rowk[e] = map[k] (or k) for every entry e of row k.
__global__ void k_row_expand(int ns, const int *ptr, const int *map, int *rowk)
{
    const int g = threadIdx.x & 15;         // one 16-lane group per row
    const int ngroups = (gridDim.x * blockDim.x) >> 4;
    for (int k = (blockIdx.x * blockDim.x + threadIdx.x) >> 4; k < ns; k += ngroups) {
        const int v = map ? map[k] : k;
        for (int e = ptr[k] + g; e < ptr[k + 1]; e += 16) rowk[e] = v;
    }
}

__global__ void k_segment_bounds(int n, const uint32_t *keys, int shift, int nseg, int *ptr)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += gridDim.x * blockDim.x) {
        int lo = (i == 0) ? -1 : (int)(keys[i - 1] >> shift);
        int hi = (i == n) ? nseg : (int)(keys[i] >> shift);
        if (hi > nseg) hi = nseg;
        for (int r = lo + 1; r <= hi; ++r) ptr[r] = i;
    }
}

__global__ void k_level_keys(int ns, const int *level, const int *lptr, int len_bits, uint32_t *keys)
{
    const int lmax = (1 << len_bits) - 1;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < ns; k += gridDim.x * blockDim.x) {
        int len = lptr[k + 1] - lptr[k];
        if (len > lmax) len = lmax;
        keys[k] = ((uint32_t)level[k] << len_bits) | (uint32_t)(lmax - len);
    }
}

// Batched systems: (level, system, descending row length) so that the rows
// of one system inside a level are contiguous.
__global__ void k_level_keys_sys(int ns, const int *level, const int *lptr, int len_bits, const int *sys, int sys_bits,
                                 uint32_t *keys)
{
    const int lmax = (1 << len_bits) - 1;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < ns; k += gridDim.x * blockDim.x) {
        int len = lptr[k + 1] - lptr[k];
        if (len > lmax) len = lmax;
        uint32_t ls = ((uint32_t)level[k] << sys_bits) | (uint32_t)sys[k];
        keys[k] = (ls << len_bits) | (uint32_t)(lmax - len);
    }
}

// keys[q] = level << 3 | class of row order[q]; class 0 = widest (log2 of the
// wider of its L and L^T rows >= 7), 7 = empty rows.
__global__ void k_level_bucket_keys(int ns, const int *order, const int *level, const int *l_ptr, const int *lt_ptr,
                                    uint32_t *keys)
{
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ns; q += gridDim.x * blockDim.x) {
        const int row = order[q];
        const int a = l_ptr[row + 1] - l_ptr[row], b = lt_ptr[row + 1] - lt_ptr[row];
        const int w = a > b ? a : b;
        const int lg = 31 - __clz(w + 1);          // floor(log2(w + 1))
        const int cls = 7 - (lg < 7 ? lg : 7);
        keys[q] = ((uint32_t)level[row] << 3) | (uint32_t)cls;
    }
}

// keys of rows order[q]: level << 2 | sweep-group class, class c = the number
// of doublings from one lane per row that a row of width w needs to hold at
// most eight entries per lane (w <= 8: 0, <= 16: 1, <= 32: 2, else 3)
__global__ void k_level_group_keys(int ns, const int *order, const int *level, const int *l_ptr, const int *lt_ptr,
                                   uint32_t *keys)
{
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ns; q += gridDim.x * blockDim.x) {
        const int row = order[q];
        const int a = l_ptr[row + 1] - l_ptr[row], b = lt_ptr[row + 1] - lt_ptr[row];
        const int w = a > b ? a : b;
        const int cls = w <= 8 ? 0 : (w <= 16 ? 1 : (w <= 32 ? 2 : 3));
        keys[q] = ((uint32_t)level[row] << 2) | (uint32_t)cls;
    }
}

__global__ void k_inverse_perm(int n, const int *order, int *pos_of)
{
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) pos_of[order[p]] = p;
}

template <typename T>
__global__ void k_gather_t(const T *src, const int *idx, T *dst, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[idx[i]];
}

__global__ void k_mul64(const int *in, int *out, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = in[i] * kWave;
}

__device__ __forceinline__ void slice_row(int c, const int *start, const int *cnt, const int *order, int nrows,
                                          int lane, int &row)
{
    if (start) {
        row = (lane < cnt[c]) ? order[start[c] + lane] : -1;
    } else {
        int p = c * kWave + lane;
        row = (p < nrows) ? (order ? order[p] : p) : -1;
    }
}

__global__ void k_ell_width(int nslices, const int *start, const int *cnt, const int *order, int nrows,
                            const int *ptr, int *width)
{
    const int lane = lane_id();
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * blockDim.x / kWave;
    for (int c = wave; c < nslices; c += nwaves) {
        // a chunk holds up to 256 rows: 64 at a time
        const int rows = start ? cnt[c] : kWave;
        int len = 0;
        for (int i0 = 0; i0 < rows; i0 += kWave) {
            int row;
            slice_row(c, start, cnt, order, nrows, i0 + lane, row);
            const int l = row >= 0 ? ptr[row + 1] - ptr[row] : 0;
            len = l > len ? l : len;
        }
        for (int off = 32; off > 0; off >>= 1) {
            int o = __shfl_xor(len, off, kWave);
            len = o > len ? o : len;
        }
        if (lane == 0) width[c] = len;
    }
}

template <typename TS, typename TD>
__global__ void k_ell_fill(int nslices, const int *start, const int *cnt, const int *order, int nrows, EllSrc src,
                           const int *off, const int *width, int pad_ind, TD pad_val, int *ell_ind, TD *ell_val)
{
    const int lane = lane_id();
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * blockDim.x / kWave;
    const TS *sval = (const TS *)src.val;
    for (int c = wave; c < nslices; c += nwaves) {
        int row;
        slice_row(c, start, cnt, order, nrows, lane, row);
        int b = row >= 0 ? src.ptr[row] : 0;
        int len = row >= 0 ? src.ptr[row + 1] - b : 0;
        const int w = width[c];
        const int o = off[c];
        // four entries per lane in flight (index, column and value, position)
        constexpr int U = 4;
        for (int t0 = 0; t0 < w; t0 += U) {
            int e[U], col[U];
            TD v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int t = t0 + u;
                e[u] = t < len ? (src.eidx ? src.eidx[b + t] : b + t) : -1;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                col[u] = e[u] >= 0 ? (src.rowk ? src.rowk[e[u]] : src.ind[e[u]]) : 0;
                v[u] = e[u] >= 0 ? (TD)sval[e[u]] : pad_val;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int t = t0 + u;
                if (t < w) {
                    ell_ind[o + t * kWave + lane] = e[u] >= 0 ? (src.colmap ? src.colmap[col[u]] : col[u]) : pad_ind;
                    ell_val[o + t * kWave + lane] = v[u];
                }
            }
        }
    }
}

// Widest L / L^T row of every level.  Positions are sorted by level, so a
// wave walks a contiguous span of kLevelSpan positions keeping the running
// maximum of the current level and adds it with one atomicMax when the level
// changes (a wide level's words would otherwise take one returning atomic per
// 64 rows, all on one address).
constexpr int kLevelSpan = 1024;
__global__ void k_level_width(int ns, const int *order, const int *level, const uint32_t *seg_keys, int seg_shift,
                              const int *l_ptr, const int *lt_ptr, int *lw)
{
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * blockDim.x / kWave;
    for (int s0 = wave * kLevelSpan; s0 < ns; s0 += nwaves * kLevelSpan) {
        int curL = -1, curW = 0;            // wave-uniform
        for (int base = s0; base < s0 + kLevelSpan && base < ns; base += kWave) {
            // lanes past ns count as width 0 at level -1
            const int p = base + lane_id();
            int L = -1, w = 0;
            if (p < ns) {
                const int row = order[p];
                L = seg_keys ? (int)(seg_keys[p] >> seg_shift) : level[row];
                const int a = l_ptr[row + 1] - l_ptr[row], b = lt_ptr[row + 1] - lt_ptr[row];
                w = a > b ? a : b;
            }
            const int L0 = __shfl(L, 0, kWave);
            const bool uniform = __all(L == L0 || L < 0);
            if (uniform) {
                for (int off = 32; off > 0; off >>= 1) {
                    const int o = __shfl_xor(w, off, kWave);
                    w = o > w ? o : w;
                }
                if (L0 != curL) {
                    if (curL >= 0 && lane_id() == 0) atomicMax(lw + curL, curW);
                    curL = L0;
                    curW = w;
                } else {
                    curW = w > curW ? w : curW;
                }
            } else {
                if (curL >= 0 && lane_id() == 0) atomicMax(lw + curL, curW);
                curL = -1;
                curW = 0;
                if (L >= 0) atomicMax(lw + L, w);
            }
        }
        if (curL >= 0 && lane_id() == 0) atomicMax(lw + curL, curW);
    }
}

__global__ void k_chunk_count(int nlevels, const int *lvl_ptr, const int *lw, int gmin, int *cnt)
{
    for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < nlevels; l += gridDim.x * blockDim.x) {
        const int r = chunk_rows(lw[l], gmin);
        cnt[l] = (lvl_ptr[l + 1] - lvl_ptr[l] + r - 1) / r;
    }
}

__global__ void k_chunk_fill(int nchunks, int nlevels, const int *lvl_ptr, const int *chunk_ptr, const int *lw,
                             int gmin, int *chunk_start, int *chunk_cnt)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nchunks; c += gridDim.x * blockDim.x) {
        int lo = 0, hi = nlevels - 1;          // last level with chunk_ptr[l] <= c
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (chunk_ptr[mid] <= c) lo = mid; else hi = mid - 1;
        }
        const int r = chunk_rows(lw[lo], gmin);
        int start = lvl_ptr[lo] + (c - chunk_ptr[lo]) * r;
        int end = lvl_ptr[lo + 1];
        chunk_start[c] = start;
        chunk_cnt[c] = (end - start) < r ? (end - start) : r;
    }
}

__global__ void k_pack_meta(int nchunks, const int *start, const int *cnt, const int *off, const int *slices,
                            int4 *meta)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nchunks; c += gridDim.x * blockDim.x)
        meta[c] = make_int4(start[c], cnt[c], off[c], slices[c]);
}
// ---------------------------------------------------------------- tail split
__device__ __forceinline__ int src_entry(const EllSrc &src, int b, int et, int &pos)
{
    const int e = src.eidx ? src.eidx[b + et] : b + et;
    const int col = src.rowk ? src.rowk[e] : src.ind[e];
    pos = src.colmap ? src.colmap[col] : col;
    return e;
}

// One wave per tail row: entries with column position < tp0 and >= tp0.
__global__ void k_tail_count(int ntail, int tp0, const int *order, EllSrc src, int *cnt_b, int *cnt_i)
{
    const int lane = lane_id();
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * blockDim.x / kWave;
    for (int q = wave; q < ntail; q += nwaves) {
        const int row = order[tp0 + q];
        const int b = src.ptr[row], len = src.ptr[row + 1] - b;
        int nb = 0;
        for (int t = lane; t < len; t += kWave) {
            int pos;
            (void)src_entry(src, b, t, pos);
            nb += pos < tp0 ? 1 : 0;
        }
        for (int off = 32; off > 0; off >>= 1) nb += __shfl_xor(nb, off, kWave);
        if (lane == 0) {
            cnt_b[q] = nb;
            cnt_i[q] = len - nb;
        }
    }
}

// One wave per tail row: stable compaction of the two parts (ballot ranks).
template <typename TD>
__global__ void k_tail_split(int ntail, int tp0, const int *order, EllSrc src, const int *ptr_b, const int *ptr_i,
                             int *ind_b, TD *val_b, int *ind_i, TD *val_i)
{
    const int lane = lane_id();
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * blockDim.x / kWave;
    const TD *sval = (const TD *)src.val;
    for (int q = wave; q < ntail; q += nwaves) {
        const int row = order[tp0 + q];
        const int b = src.ptr[row], len = src.ptr[row + 1] - b;
        int ob = ptr_b[q], oi = ptr_i[q];
        for (int base = 0; base < len; base += kWave) {
            const int t = base + lane;
            int pos = 0, e = 0;
            if (t < len) e = src_entry(src, b, t, pos);
            const bool isb = t < len && pos < tp0;
            const bool isi = t < len && pos >= tp0;
            const uint64_t mb = ballot(isb), mi = ballot(isi);
            if (isb) {
                const int o = ob + lanes_below(mb);
                ind_b[o] = pos;
                val_b[o] = sval[e];
            }
            if (isi) {
                const int o = oi + lanes_below(mi);
                ind_i[o] = pos - tp0;
                val_i[o] = sval[e];
            }
            ob += __popcll(mb);
            oi += __popcll(mi);
        }
    }
}

// One wave per tail level: widest internal row -> slices and block size.
__global__ void k_tail_layout(int ntl, const int *lstart, const int *lcnt, int tp0, const int *ptr_i, int *slices,
                              int *size)
{
    const int lane = lane_id();
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * blockDim.x / kWave;
    for (int L = wave; L < ntl; L += nwaves) {
        const int q0 = lstart[L] - tp0, cnt = lcnt[L];
        int w = 0;
        for (int i = lane; i < cnt; i += kWave) {
            const int len = ptr_i[q0 + i + 1] - ptr_i[q0 + i];
            w = len > w ? len : w;
        }
        for (int off = 32; off > 0; off >>= 1) {
            const int o = __shfl_xor(w, off, kWave);
            w = o > w ? o : w;
        }
        if (lane == 0) {
            const int G = tail_group(cnt);
            const int sl = (w + G - 1) / G;
            slices[L] = sl;
            size[L] = (sl + kTailBlock - 1) / kTailBlock * kTailBlock * sweep_rows_pow2(cnt) * G;
        }
    }
}

template <typename TD>
__global__ void k_tail_fill(int ntl, const int *lstart, const int *lcnt, int tp0, const int *ptr_i, const int *ind_i,
                            const TD *val_i, const int *off, const int *slices, int pad_ind, int *ell_ind, TD *ell_val)
{
    const int lane = lane_id();
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * blockDim.x / kWave;
    for (int L = wave; L < ntl; L += nwaves) {
        const int cnt = lcnt[L], q0 = lstart[L] - tp0;
        const int P = sweep_rows_pow2(cnt), G = tail_group(cnt);
        const int stride = P * G, o = off[L];
        const int blk = stride * kTailBlock;
        const int total = (slices[L] + kTailBlock - 1) / kTailBlock * blk;
        for (int idx = lane; idx < total; idx += kWave) {
            // idx = block * blk + lane slot * kTailBlock + u: slice block * kTailBlock + u
            const int b0 = idx / blk, r0 = idx - b0 * blk;
            const int t = r0 / kTailBlock, u = r0 - t * kTailBlock;
            const int sl = b0 * kTailBlock + u;
            const int i = t / G, g = t - i * G;
            const int et = sl * G + g;
            int ind = pad_ind;
            TD v = (TD)0;
            if (i < cnt) {
                const int b = ptr_i[q0 + i];
                if (et < ptr_i[q0 + i + 1] - b) {
                    ind = ind_i[b + et];
                    v = val_i[b + et];
                }
            }
            ell_ind[o + idx] = ind;
            ell_val[o + idx] = v;
        }
    }
}

__global__ void k_sweep_layout(int nchunks, const int *cnt, const int *width, int *slices, int *size)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nchunks; c += gridDim.x * blockDim.x) {
        int g = sweep_group(cnt[c]);
        int sl = (width[c] + g - 1) / g;
        slices[c] = sl;
        size[c] = sl * sweep_rows_pow2(cnt[c]) * g;
    }
}

constexpr int kFillWaves = 4;     // waves per k_sweep_fill workgroup (one chunk per wave at a time)

template <typename TS, typename TD>
__global__ void __launch_bounds__(kFillWaves * kWave)
    k_sweep_fill(int nchunks, const int *start, const int *cnt, const int *order, EllSrc src, const int *off,
                 const int *slices, int pad_ind, TD pad_val, int *ell_ind, TD *ell_val)
{
    // source range of every row of the wave's chunk (up to 256 rows)
    __shared__ int s_rb[kFillWaves][kWaveRows256], s_rl[kFillWaves][kWaveRows256];
    const int lane = lane_id();
    const int wl = threadIdx.x / kWave;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * blockDim.x / kWave;
    const TS *sval = (const TS *)src.val;
    int *rbs = s_rb[wl], *rls = s_rl[wl];
    for (int c = wave; c < nchunks; c += nwaves) {
        const int nc = cnt[c];                  // <= 256 rows (sweep chunks)
        const int P = sweep_rows_pow2(nc);
        const int G = sweep_group(nc);
        const int stride = P * G;
        const int total = slices[c] * stride;
        const int o = off[c];
        // the previous chunk's reads of the wave's slots are done before they
        // are overwritten (LDS operations of one wave complete in order; the
        // fences keep the compiler from moving accesses across)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int i = lane; i < nc; i += kWave) {
            const int row = order[start[c] + i];
            const int b = src.ptr[row];
            rbs[i] = b;
            rls[i] = src.ptr[row + 1] - b;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // four output slots per lane in flight: index, then column and value,
        // then the column's position
        constexpr int U = 4;
        for (int idx0 = 0; idx0 < total; idx0 += U * kWave) {
            int e[U], col[U];
            TD v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int idx = idx0 + u * kWave + lane;
                const int sl = idx / stride, t = idx - sl * stride;
                const int i = t / G, g = t - i * G;
                const int et = sl * G + g;
                const bool in = idx < total && i < nc;
                const int b = in ? rbs[i] : 0;
                const int len = in ? rls[i] : 0;
                e[u] = (in && et < len) ? (src.eidx ? src.eidx[b + et] : b + et) : -1;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                col[u] = e[u] >= 0 ? (src.rowk ? src.rowk[e[u]] : src.ind[e[u]]) : 0;
                v[u] = e[u] >= 0 ? (TD)sval[e[u]] : pad_val;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int idx = idx0 + u * kWave + lane;
                if (idx < total) {
                    ell_ind[o + idx] = e[u] >= 0 ? (src.colmap ? src.colmap[col[u]] : col[u]) : pad_ind;
                    ell_val[o + idx] = v[u];
                }
            }
        }
    }
}
__global__ void k_pos_csr_count(int n, const int *order, const int *src_ptr, int *cnt)
{
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        int row = order[p];
        cnt[p] = src_ptr[row + 1] - src_ptr[row];
    }
}

template <typename TS, typename TD>
__global__ void k_pos_csr_fill(int n, const int *order, EllSrc src, const int *ptr, int *ind, TD *val)
{
    const TS *sval = (const TS *)src.val;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        int row = order[p];
        int b = src.ptr[row], len = src.ptr[row + 1] - b, o = ptr[p];
        for (int t = 0; t < len; ++t) {
            int e = src.eidx ? src.eidx[b + t] : b + t;
            int col = src.rowk ? src.rowk[e] : src.ind[e];
            ind[o + t] = src.colmap ? src.colmap[col] : col;
            val[o + t] = (TD)sval[e];
        }
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

void launch_level_max(int ns, const int *level, unsigned long long *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_level_max, dim3(1024), dim3(256), 0, s, ns, level, out);
}
void launch_iota(int *v, int n, hipStream_t s) { hipLaunchKernelGGL(k_iota, dim3(2048), dim3(256), 0, s, v, n); }
void launch_copy_keys(const int *src, uint32_t *dst, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy_keys, dim3(2048), dim3(256), 0, s, src, dst, n);
}
__global__ void k_tile_keys(int ntiles, const int *idx_user, uint32_t *keys, int *vals)
{
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < ntiles; t += gridDim.x * blockDim.x) {
        keys[t] = (uint32_t)idx_user[(size_t)t * kWave];
        vals[t] = t;
    }
}
void launch_tile_keys(int ntiles, const int *idx_user, uint32_t *keys, int *vals, hipStream_t s)
{
    hipLaunchKernelGGL(k_tile_keys, dim3(256), dim3(256), 0, s, ntiles, idx_user, keys, vals);
}
void launch_row_expand(int ns, const int *ptr, const int *map, int *rowk, hipStream_t s)
{
    hipLaunchKernelGGL(k_row_expand, dim3(2048), dim3(256), 0, s, ns, ptr, map, rowk);
}
void launch_segment_bounds(int n, const uint32_t *keys, int shift, int nseg, int *ptr, hipStream_t s)
{
    hipLaunchKernelGGL(k_segment_bounds, dim3(2048), dim3(256), 0, s, n, keys, shift, nseg, ptr);
}
void launch_level_keys(int ns, const int *level, const int *lptr, int len_bits, uint32_t *keys, hipStream_t s)
{
    hipLaunchKernelGGL(k_level_keys, dim3(2048), dim3(256), 0, s, ns, level, lptr, len_bits, keys);
}
void launch_level_keys_sys(int ns, const int *level, const int *lptr, int len_bits, const int *sys, int sys_bits,
                           uint32_t *keys, hipStream_t s)
{
    hipLaunchKernelGGL(k_level_keys_sys, dim3(2048), dim3(256), 0, s, ns, level, lptr, len_bits, sys, sys_bits, keys);
}
void launch_level_bucket_keys(int ns, const int *order, const int *level, const int *l_ptr, const int *lt_ptr,
                              uint32_t *keys, hipStream_t s)
{
    hipLaunchKernelGGL(k_level_bucket_keys, dim3(2048), dim3(256), 0, s, ns, order, level, l_ptr, lt_ptr, keys);
}
void launch_level_group_keys(int ns, const int *order, const int *level, const int *l_ptr, const int *lt_ptr,
                             uint32_t *keys, hipStream_t s)
{
    hipLaunchKernelGGL(k_level_group_keys, dim3(2048), dim3(256), 0, s, ns, order, level, l_ptr, lt_ptr, keys);
}
void launch_inverse_perm(int n, const int *order, int *pos_of, hipStream_t s)
{
    hipLaunchKernelGGL(k_inverse_perm, dim3(2048), dim3(256), 0, s, n, order, pos_of);
}
template <typename T>
void launch_gather_t(const T *src, const int *idx, T *dst, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_gather_t<T>, dim3(2048), dim3(256), 0, s, src, idx, dst, n);
}
// keys[p] = locality rank of factored row p's input vertex (rows without one,
// such as an appended ground vertex, after all of them)
__global__ void k_locality_keys(int ns, const int *perm, int nuser, const int *rank, uint32_t *keys)
{
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < ns; p += gridDim.x * blockDim.x) {
        const int u = perm[p];
        keys[p] = (uint32_t)((u >= 0 && u < nuser) ? rank[u] : nuser);
    }
}
void launch_locality_keys(int ns, const int *perm, int nuser, const int *rank, uint32_t *keys, hipStream_t s)
{
    hipLaunchKernelGGL(k_locality_keys, dim3(2048), dim3(256), 0, s, ns, perm, nuser, rank, keys);
}
void launch_gather_i(const int *src, const int *idx, int *dst, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_gather_t<int>, dim3(2048), dim3(256), 0, s, src, idx, dst, n);
}
void launch_mul64(const int *in, int *out, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_mul64, dim3(1024), dim3(256), 0, s, in, out, n);
}
void launch_ell_width(int nslices, const int *start, const int *cnt, const int *order, int nrows, const int *ptr,
                      int *width, hipStream_t s)
{
    int g = (nslices + 3) / 4;
    if (g > 16384) g = 16384;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_ell_width, dim3(g), dim3(256), 0, s, nslices, start, cnt, order, nrows, ptr, width);
}
template <typename TS, typename TD>
void launch_ell_fill(int nslices, const int *start, const int *cnt, const int *order, int nrows, EllSrc src,
                     const int *off, const int *width, int pad_ind, TD pad_val, int *ell_ind, TD *ell_val,
                     hipStream_t s)
{
    int g = (nslices + 3) / 4;
    if (g > 16384) g = 16384;
    if (g < 1) g = 1;
    hipLaunchKernelGGL((k_ell_fill<TS, TD>), dim3(g), dim3(256), 0, s, nslices, start, cnt, order, nrows, src, off,
                       width, pad_ind, pad_val, ell_ind, ell_val);
}

void launch_level_width(int ns, const int *order, const int *level, const uint32_t *seg_keys, int seg_shift,
                        const int *l_ptr, const int *lt_ptr, int *lw, hipStream_t s)
{
    hipLaunchKernelGGL(k_level_width, dim3(2048), dim3(256), 0, s, ns, order, level, seg_keys, seg_shift, l_ptr, lt_ptr,
                       lw);
}
void launch_chunk_count(int nlevels, const int *lvl_ptr, const int *lw, int gmin, int *cnt, hipStream_t s)
{
    hipLaunchKernelGGL(k_chunk_count, dim3(256), dim3(256), 0, s, nlevels, lvl_ptr, lw, gmin, cnt);
}
void launch_chunk_fill(int nchunks, int nlevels, const int *lvl_ptr, const int *chunk_ptr, const int *lw, int gmin,
                       int *chunk_start, int *chunk_cnt, hipStream_t s)
{
    hipLaunchKernelGGL(k_chunk_fill, dim3(1024), dim3(256), 0, s, nchunks, nlevels, lvl_ptr, chunk_ptr, lw, gmin,
                       chunk_start, chunk_cnt);
}
void launch_pack_meta(int nchunks, const int *start, const int *cnt, const int *off, const int *slices, int4 *meta,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_pack_meta, dim3(1024), dim3(256), 0, s, nchunks, start, cnt, off, slices, meta);
}

static int waves_grid(int items)
{
    int g = (items + 3) / 4;
    if (g > 16384) g = 16384;
    return g < 1 ? 1 : g;
}
void launch_tail_count(int ntail, int tp0, const int *order, EllSrc src, int *cnt_b, int *cnt_i, hipStream_t s)
{
    hipLaunchKernelGGL(k_tail_count, dim3(waves_grid(ntail)), dim3(256), 0, s, ntail, tp0, order, src, cnt_b, cnt_i);
}
template <typename TD>
void launch_tail_split(int ntail, int tp0, const int *order, EllSrc src, const int *ptr_b, const int *ptr_i,
                       int *ind_b, TD *val_b, int *ind_i, TD *val_i, hipStream_t s)
{
    hipLaunchKernelGGL(k_tail_split<TD>, dim3(waves_grid(ntail)), dim3(256), 0, s, ntail, tp0, order, src, ptr_b,
                       ptr_i, ind_b, val_b, ind_i, val_i);
}
void launch_tail_layout(int ntl, const int *lstart, const int *lcnt, int tp0, const int *ptr_i, int *slices,
                        int *size, hipStream_t s)
{
    hipLaunchKernelGGL(k_tail_layout, dim3(waves_grid(ntl)), dim3(256), 0, s, ntl, lstart, lcnt, tp0, ptr_i, slices,
                       size);
}
template <typename TD>
void launch_tail_fill(int ntl, const int *lstart, const int *lcnt, int tp0, const int *ptr_i, const int *ind_i,
                      const TD *val_i, const int *off, const int *slices, int pad_ind, int *ell_ind, TD *ell_val,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_tail_fill<TD>, dim3(waves_grid(ntl)), dim3(256), 0, s, ntl, lstart, lcnt, tp0, ptr_i, ind_i,
                       val_i, off, slices, pad_ind, ell_ind, ell_val);
}
template void launch_tail_split<float>(int, int, const int *, EllSrc, const int *, const int *, int *, float *, int *,
                                       float *, hipStream_t);
template void launch_tail_split<double>(int, int, const int *, EllSrc, const int *, const int *, int *, double *,
                                        int *, double *, hipStream_t);
template void launch_tail_fill<float>(int, const int *, const int *, int, const int *, const int *, const float *,
                                      const int *, const int *, int, int *, float *, hipStream_t);
template void launch_tail_fill<double>(int, const int *, const int *, int, const int *, const int *, const double *,
                                       const int *, const int *, int, int *, double *, hipStream_t);

template void launch_lt_fill<float>(const AssembleArgs &, const int *, int *, float *, float *, hipStream_t);
template void launch_lt_fill<double>(const AssembleArgs &, const int *, int *, double *, double *, hipStream_t);
void launch_sweep_layout(int nchunks, const int *cnt, const int *width, int *slices, int *size, hipStream_t s)
{
    hipLaunchKernelGGL(k_sweep_layout, dim3(1024), dim3(256), 0, s, nchunks, cnt, width, slices, size);
}

template <typename TS, typename TD>
void launch_sweep_fill(int nchunks, const int *start, const int *cnt, const int *order, EllSrc src, const int *off,
                       const int *slices, int pad_ind, TD pad_val, int *ell_ind, TD *ell_val, hipStream_t s)
{
    int g = (nchunks + 3) / 4;
    if (g > 16384) g = 16384;
    if (g < 1) g = 1;
    hipLaunchKernelGGL((k_sweep_fill<TS, TD>), dim3(g), dim3(kFillWaves * kWave), 0, s, nchunks, start, cnt, order,
                       src, off, slices, pad_ind, pad_val, ell_ind, ell_val);
}

template void launch_sweep_fill<float, float>(int, const int *, const int *, const int *, EllSrc, const int *,
                                             const int *, int, float, int *, float *, hipStream_t);
template void launch_sweep_fill<double, double>(int, const int *, const int *, const int *, EllSrc, const int *,
                                               const int *, int, double, int *, double *, hipStream_t);
void launch_pos_csr_count(int n, const int *order, const int *src_ptr, int *cnt, hipStream_t s)
{
    hipLaunchKernelGGL(k_pos_csr_count, dim3(2048), dim3(256), 0, s, n, order, src_ptr, cnt);
}
template <typename TS, typename TD>
void launch_pos_csr_fill(int n, const int *order, EllSrc src, const int *ptr, int *ind, TD *val, hipStream_t s)
{
    hipLaunchKernelGGL((k_pos_csr_fill<TS, TD>), dim3(2048), dim3(256), 0, s, n, order, src, ptr, ind, val);
}
template void launch_pos_csr_fill<float, float>(int, const int *, EllSrc, const int *, int *, float *, hipStream_t);
template void launch_pos_csr_fill<double, double>(int, const int *, EllSrc, const int *, int *, double *, hipStream_t);

template void launch_gather_t<float>(const float *, const int *, float *, int, hipStream_t);
template void launch_gather_t<double>(const double *, const int *, double *, int, hipStream_t);
template void launch_ell_fill<float, float>(int, const int *, const int *, const int *, int, EllSrc, const int *,
                                           const int *, int, float, int *, float *, hipStream_t);
template void launch_ell_fill<double, double>(int, const int *, const int *, const int *, int, EllSrc, const int *,
                                             const int *, int, double, int *, double *, hipStream_t);

}  // namespace rc
