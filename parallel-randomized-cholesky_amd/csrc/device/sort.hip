// sort.hip -- stable LSD radix sort of (uint32 key, int32 value) pairs.
//
// Used by the assembly to order rows by (level, row length) and to transpose
// the factor (entries sorted by row, stable in column order), so both are
// deterministic and free of contended atomics.  8-bit digits; each pass is a
// per-tile digit histogram (LDS), an exclusive scan of the digit-major count
// table, and a stable scatter whose in-tile ranks come from wave ballots.
#include "common.cuh"
#include "kernels.h"

namespace rc {

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kDigits = 256;
constexpr int kSortWaves = kSortThreads / kWave;
static_assert(kSortThreads == kDigits, "the tile scan gives one digit to every thread");

__global__ void __launch_bounds__(kSortThreads) k_radix_hist(const uint32_t *keys, int n, int shift, int ntiles,
                                                             int *counts)
{
    __shared__ int hist[kDigits];
    for (int d = threadIdx.x; d < kDigits; d += kSortThreads) hist[d] = 0;
    __syncthreads();
    const int base = blockIdx.x * kSortTile;
    for (int j = 0; j < kSortItems; ++j) {
        int i = base + j * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&hist[(keys[i] >> shift) & 0xFF], 1);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < kDigits; d += kSortThreads) counts[d * ntiles + blockIdx.x] = hist[d];
}

// Stable scatter of one tile: every item's rank inside the tile in digit order
// (digits before it in the tile, plus items of its digit earlier in the tile:
// wave ballots, then earlier waves and rounds) places it in an LDS copy of the
// tile sorted by digit; the copy is then written out in order, so each
// digit's run goes to consecutive addresses of its output range.
__global__ void __launch_bounds__(kSortThreads) k_radix_scatter(const uint32_t *keys, const int *vals, int n,
                                                                int shift, int ntiles, const int *counts,
                                                                const int *offsets, uint32_t *keys_out, int *vals_out)
{
    __shared__ int tstart[kDigits];      // first tile slot of each digit
    __shared__ int gstart[kDigits];      // output index of the tile's first item of each digit
    __shared__ int running[kDigits];     // items of each digit placed so far
    __shared__ int wcount[kSortWaves][kDigits];
    __shared__ uint32_t skey[kSortTile];
    __shared__ int sval[kSortTile];
    const int lane = lane_id();
    const int wid = threadIdx.x / kWave;
    // exclusive scan of the tile's digit counts (one digit per thread)
    {
        const int d = threadIdx.x;       // kSortThreads == kDigits
        const int c = counts[d * ntiles + blockIdx.x];
        int x = c;
        for (int off = 1; off < kWave; off <<= 1) {
            const int o = __shfl_up(x, off, kWave);
            if (lane >= off) x += o;
        }
        if (lane == kWave - 1) wcount[0][wid] = x;     // wave totals
        __syncthreads();
        int before = 0;
        for (int w = 0; w < wid; ++w) before += wcount[0][w];
        tstart[d] = before + x - c;
        gstart[d] = offsets[d * ntiles + blockIdx.x];
        running[d] = 0;
        __syncthreads();
        for (int w = 0; w < kSortWaves; ++w) wcount[w][d] = 0;
        __syncthreads();
    }
    const int base = blockIdx.x * kSortTile;
    const int tile_n = min(kSortTile, n - base);
    for (int j = 0; j < kSortItems; ++j) {
        const int i = base + j * kSortThreads + threadIdx.x;
        const bool valid = i < n;
        const uint32_t key = valid ? keys[i] : 0u;
        const int val = valid ? vals[i] : 0;
        const int d = (int)((key >> shift) & 0xFF);
        // lanes of this wave holding the same digit
        uint64_t peers = ballot(valid);
        for (int b = 0; b < 8; ++b) {
            const uint64_t m = ballot(((d >> b) & 1) != 0);
            peers &= ((d >> b) & 1) ? m : ~m;
        }
        const int rank = lanes_below(peers);
        if (valid && rank == 0) wcount[wid][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int pos = tstart[d] + running[d] + rank;
            for (int w = 0; w < wid; ++w) pos += wcount[w][d];
            skey[pos] = key;
            sval[pos] = val;
        }
        __syncthreads();
        {
            const int dd = threadIdx.x;
            int sum = 0;
            for (int w = 0; w < kSortWaves; ++w) {
                sum += wcount[w][dd];
                wcount[w][dd] = 0;
            }
            running[dd] += sum;
        }
        __syncthreads();
    }
    for (int p = threadIdx.x; p < tile_n; p += kSortThreads) {
        const uint32_t key = skey[p];
        const int d = (int)((key >> shift) & 0xFF);
        const int o = gstart[d] + (p - tstart[d]);
        keys_out[o] = key;
        vals_out[o] = sval[p];
    }
}

}  // namespace

size_t radix_tmp_ints(int n)
{
    size_t ntiles = (size_t)(n + kSortTile - 1) / kSortTile;
    if (ntiles < 1) ntiles = 1;
    size_t cnt = ntiles * kDigits;
    return (cnt + 1) * 2 + scan_tmp_elems((int)cnt) + 16;
}

// Sorts (k0, v0) by key; the result is left in (k0, v0).  k1/v1 are scratch of
// n elements, tmp holds radix_tmp_ints(n) ints.
void radix_sort_pairs(uint32_t *k0, int *v0, uint32_t *k1, int *v1, int n, int key_bits, int *tmp, hipStream_t s)
{
    if (n <= 1) return;
    int ntiles = (n + kSortTile - 1) / kSortTile;
    int cnt = ntiles * kDigits;
    int *counts = tmp;                  // cnt
    int *offsets = tmp + (cnt + 1);     // cnt + 1: scan_exclusive writes the total at [cnt]
    int *scan_tmp = tmp + 2 * (cnt + 1);
    int passes = (key_bits + 7) / 8;
    if (passes < 1) passes = 1;
    uint32_t *ka = k0, *kb = k1;
    int *va = v0, *vb = v1;
    for (int p = 0; p < passes; ++p) {
        int shift = 8 * p;
        hipLaunchKernelGGL(k_radix_hist, dim3(ntiles), dim3(kSortThreads), 0, s, ka, n, shift, ntiles, counts);
        scan_exclusive(counts, offsets, cnt, scan_tmp, s);
        hipLaunchKernelGGL(k_radix_scatter, dim3(ntiles), dim3(kSortThreads), 0, s, ka, va, n, shift, ntiles, counts,
                           offsets, kb, vb);
        uint32_t *tk = ka; ka = kb; kb = tk;
        int *tv = va; va = vb; vb = tv;
    }
    if (ka != k0) {
        (void)hipMemcpyAsync(k0, ka, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
        (void)hipMemcpyAsync(v0, va, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice, s);
    }
}

}  // namespace rc
