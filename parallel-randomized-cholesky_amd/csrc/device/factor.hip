// factor.hip -- persistent randomized-elimination kernel for gfx950.
//
// Semantics follow cpu_implementation/driver_local.cpp:302-707 and
// gpu_implementation/driver.cu:43-1302 (the reference's per-column step):
// gather the column's strict-lower entries and pending fill edges, merge
// duplicate rows (summing weight and multiplicity), sort by weight, take the
// forward cumulative sums C and S = C[d-1], and for every entry i < d-1 draw
// u in (0,1], pick the first j > i with C[j] - C[i] >= u (S - C[i]) and emit
// the fill edge (max, min, w_i (S - C[i]) / S) into column min; the factor
// column holds -w/S and the pivot is S.
//
// MI355X design:
//   * one wave64 (one workgroup) per column at a time; columns come from a
//     ready queue claimed with one atomic per column (work stealing -- any
//     grid size is safe, waves never wait while holding work);
//   * gather, duplicate merge, sorts and the prefix sum run in LDS (global
//     per-wave scratch for columns above the LDS capacity);
//   * the sampling draw is Philox4x32-10 keyed by (seed, i, column): with the
//     canonical (row, weight) merge order and (weight, row) sampling order the
//     factor is a pure function of (matrix, seed), independent of schedule;
//   * fill edges are appended to per-column lists in HBM by one atomic
//     exchange each; dependency counters are agent-scope atomics and the
//     hand-off of a ready column is a release/acquire pair;
//   * every wait is bounded; pool or scratch exhaustion raises a status bit,
//     all waves drain, and the host regrows and reruns (same result).
#include "common.cuh"
#include "kernels.h"

namespace rc {

namespace {

constexpr int kLdsCap = 128;                        // entries per wave in LDS
constexpr int kBytesPerEntry = 8 + 8 + 8 + 4 + 4 + 4 + 4;  // sw mw C srow smult mrow mmult

struct ColBuf {
    double *sw, *mw, *C;
    int *srow, *smult, *mrow, *mmult;
};

__device__ __forceinline__ ColBuf carve(char *base, int cap)
{
    ColBuf b;
    b.sw = reinterpret_cast<double *>(base);
    b.mw = b.sw + cap;
    b.C = b.mw + cap;
    b.srow = reinterpret_cast<int *>(b.C + cap);
    b.smult = b.srow + cap;
    b.mrow = b.smult + cap;
    b.mmult = b.mrow + cap;
    return b;
}

__device__ __forceinline__ int next_pow2(int v)
{
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Ascending sorting network (bitonic with a mirrored first comparator in
// every merge stage), all comparators put the minimum at the lower index, so
// positions >= len behave as +infinity and need no padding.
// kByRow: key (row, w); otherwise key (w, row).  mult is carried if non-null.
template <bool kByRow>
__device__ void wave_sort(int *row, double *w, int *mult, int len)
{
    const int lane = lane_id();
    const int P = next_pow2(len);
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = lane; t < (P >> 1); t += kWave) {
                int i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
                int l = (j == (k >> 1)) ? (i ^ (k - 1)) : (i + j);
                if (l < len) {
                    int ri = row[i], rl = row[l];
                    double wi = w[i], wl = w[l];
                    bool gt = kByRow ? (ri > rl || (ri == rl && wi > wl))
                                     : (wi > wl || (wi == wl && ri > rl));
                    if (gt) {
                        row[i] = rl; row[l] = ri;
                        w[i] = wl; w[l] = wi;
                        if (mult) { int mi = mult[i]; mult[i] = mult[l]; mult[l] = mi; }
                    }
                }
            }
            __syncthreads();
        }
    }
}

__device__ __forceinline__ void raise_status(unsigned long long *ctr, unsigned long long bits)
{
    __hip_atomic_fetch_or(&ctr[kCtrStatus], bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Eliminates column k.  Returns false if the kernel must stop (status set).
__device__ bool eliminate_column(const FactorArgs &a, int k, char *lds)
{
    const int lane = lane_id();
    const int lo0 = a.lo_ptr[k];
    const int norig = a.lo_ptr[k + 1] - lo0;
    const int nf_in = ld_relaxed(&a.nfill[k]);
    const int m = norig + nf_in;

    if (m == 0) {                      // empty column: zero pivot (driver_local.cpp:491-498)
        if (lane == 0) {
            a.out_start[k] = 0;
            a.out_cnt[k] = 0;
            a.diag[k] = 0.0;
            fetch_add(&a.ctr[kCtrZeroPivots], 1ull);
        }
        return true;
    }

    // ---- choose the working buffer
    const int P = next_pow2(m);
    char *base;
    int cap;
    if (P <= kLdsCap) {
        base = lds;
        cap = kLdsCap;
    } else if (P <= a.scratch_cap) {
        base = a.scratch + (size_t)blockIdx.x * (size_t)a.scratch_cap * kBytesPerEntry;
        cap = a.scratch_cap;
    } else {
        if (lane == 0) {
            fetch_max(&a.ctr[kCtrScratchNeed], (unsigned long long)P);
            raise_status(a.ctr, kStatusScratch);
        }
        return false;
    }
    ColBuf b = carve(base, cap);

    // ---- 1. gather (driver_local.cpp:459-487; cpu auxilliary.hpp:89-173)
    for (int t = lane; t < norig; t += kWave) {
        b.srow[t] = a.lo_row[lo0 + t];
        b.sw[t] = a.lo_w[lo0 + t];
        b.smult[t] = 1;
    }
    if (lane == 0) {
        int f = ld_relaxed(&a.head[k]);
        int t = 0;
        while (f >= 0 && t < nf_in) {
            FillRec rec = a.pool[f];
            b.srow[norig + t] = rec.row;
            b.sw[norig + t] = rec.w;
            b.smult[norig + t] = 1;
            f = rec.next;
            ++t;
        }
        if (t != nf_in || f >= 0) raise_status(a.ctr, kStatusInconsistent);
    }
    __syncthreads();

    // ---- 2. sort by (row, w); merge duplicates left to right (:520-560)
    wave_sort<true>(b.srow, b.sw, b.smult, m);
    int d = 0;
    for (int base_i = 0; base_i < m; base_i += kWave) {
        int i = base_i + lane;
        bool head = false;
        int r = 0, ms = 0;
        double ws = 0.0;
        if (i < m) {
            r = b.srow[i];
            head = (i == 0) || (b.srow[i - 1] != r);
            if (head) {
                ws = b.sw[i];
                ms = b.smult[i];
                for (int t = i + 1; t < m && b.srow[t] == r; ++t) {
                    ws += b.sw[t];
                    ms += b.smult[t];
                }
            }
        }
        uint64_t mask = ballot(head);
        if (head) {
            int pos = d + lanes_below(mask);
            b.mrow[pos] = r;
            b.mw[pos] = ws;
            b.mmult[pos] = ms;
        }
        d += __popcll(mask);
    }
    __syncthreads();

    // ---- 3. sort merged entries by (w, row); cumulative sums (:573-593)
    for (int t = lane; t < d; t += kWave) {
        b.srow[t] = b.mrow[t];
        b.sw[t] = b.mw[t];
    }
    __syncthreads();
    wave_sort<false>(b.srow, b.sw, nullptr, d);
    if (lane == 0) {
        double c = b.sw[0];
        b.C[0] = c;
        for (int i = 1; i < d; ++i) {
            c = c + b.sw[i];
            b.C[i] = c;
        }
    }
    __syncthreads();
    const double S = b.C[d - 1];

    // ---- 4. sampled fills (:598-631; gpu auxilliary.hpp:543-592)
    const int nfill_out = d - 1;
    long long fbase = 0;
    if (nfill_out > 0) {
        if (lane == 0) fbase = (long long)fetch_add(&a.ctr[kCtrFillTop], (unsigned long long)nfill_out);
        fbase = bcast(fbase, 0);
        if (fbase + nfill_out > a.pool_cap) {
            if (lane == 0) raise_status(a.ctr, kStatusFillPool);
            return false;
        }
    }
    for (int i = lane; i < nfill_out; i += kWave) {
        double u = sample_uniform(a.seed, (uint32_t)i, (uint32_t)k);
        double Ci = b.C[i];
        double rest = S - Ci;
        double tgt = u * rest;
        int lo = i + 1, hi = d - 1;
        while (lo < hi) {
            int mid = lo + ((hi - lo) >> 1);
            if (b.C[mid] - Ci < tgt) lo = mid + 1; else hi = mid;
        }
        int ri = b.srow[i], rj = b.srow[lo];
        int rr = ri > rj ? ri : rj;
        int cc = ri > rj ? rj : ri;
        double fw = (b.sw[i] * rest) / S;
        int rec = (int)(fbase + i);
        int old = exchange(&a.head[cc], rec);
        FillRec fr;
        fr.row = rr;
        fr.next = old;
        fr.w = fw;
        a.pool[rec] = fr;
        fetch_add(&a.nfill[cc], 1);
        fetch_add(&a.dep[rr], 1);
    }

    // ---- 5. factor column (row order, -w/S), pivot S, levels (:638-645)
    long long obase = 0;
    if (lane == 0) obase = (long long)fetch_add(&a.ctr[kCtrOutTop], (unsigned long long)d);
    obase = bcast(obase, 0);
    if (obase + d > a.out_cap) {
        if (lane == 0) raise_status(a.ctr, kStatusOutPool);
        return false;
    }
    const int lk = ld_relaxed(&a.level[k]);
    for (int i = lane; i < d; i += kWave) {
        int r = b.mrow[i];
        a.out_row[obase + i] = r;
        a.out_val[obase + i] = -(b.mw[i] / S);
        fetch_max(&a.level[r], lk + 1);
    }
    if (lane == 0) {
        a.out_start[k] = obase;
        a.out_cnt[k] = d;
        a.diag[k] = S;
        fetch_add(&a.ctr[kCtrFills], (unsigned long long)nfill_out);
        fetch_add(&a.ctr[kCtrGathered], (unsigned long long)m);
        fetch_add(&a.ctr[kCtrNnzOff], (unsigned long long)d);
        fetch_max(&a.ctr[kCtrMaxDeg], (unsigned long long)d);
        fetch_max(&a.ctr[kCtrMaxLevel], (unsigned long long)lk);
    }

    // ---- 6. release, then retire the dependencies (:653-687)
    fence_release();
    for (int i = lane; i < d; i += kWave) {
        int r = b.mrow[i];
        int mm = b.mmult[i];
        int old = fetch_add(&a.dep[r], -mm);
        if (old == mm) {
            fence_acq_rel();
            unsigned long long slot = fetch_add(&a.ctr[kCtrQTail], 1ull);
            if (slot < (unsigned long long)a.n) st_relaxed(&a.queue[slot], r);
            else raise_status(a.ctr, kStatusInconsistent);   // a column retired twice
        }
    }
    __syncthreads();
    return true;
}

}  // namespace

__global__ void __launch_bounds__(kWave) k_factor(FactorArgs a)
{
    __shared__ __attribute__((aligned(16))) char lds[kLdsCap * kBytesPerEntry];
    const int lane = lane_id();
    while (true) {
        long long idx = 0;
        if (lane == 0) idx = (long long)fetch_add(&a.ctr[kCtrQHead], 1ull);
        idx = bcast(idx, 0);
        if (idx >= a.n) break;
        int col = -1;
        if (lane == 0) {
            long long t0 = wall_clock64();
            while (true) {
                col = ld_relaxed(&a.queue[idx]);
                if (col >= 0) break;
                if (ld_relaxed(&a.ctr[kCtrStatus]) != 0ull) break;
                if (wall_clock64() - t0 > a.spin_ticks) {
                    raise_status(a.ctr, kStatusTimeout);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        col = bcast(col, 0);
        if (col < 0) break;
        fence_acquire();
        if (!eliminate_column(a, col, lds)) break;
        if (ld_relaxed(&a.ctr[kCtrStatus]) != 0ull) break;
    }
}

// ---- set-up kernels ------------------------------------------------------

__global__ void k_factor_init(int n, int *dep, int *head, int *nfill, int *queue, int *level, int *out_cnt)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        dep[i] = 0;
        head[i] = -1;
        nfill[i] = 0;
        queue[i] = -1;
        level[i] = 0;
        out_cnt[i] = 0;
    }
}

// dep[r] = number of strict-lower entries (r, j) over columns j < r; equal to
// the reference's count of strict-upper entries of column r
// (pre_process.hpp:474-496) for a symmetric pattern.
__global__ void k_factor_deps(int n, const int *lo_ptr, const int *lo_row, int *dep)
{
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
        for (int p = lo_ptr[j]; p < lo_ptr[j + 1]; ++p) atomicAdd(&dep[lo_row[p]], 1);
}

__global__ void k_factor_seed_queue(int n, const int *dep, int *queue, unsigned long long *ctr)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (dep[i] == 0) {
            int slot = (int)atomicAdd(&ctr[kCtrQTail], 1ull);
            queue[slot] = i;
        }
    }
}

void launch_factor(const FactorArgs &a, int num_blocks, hipStream_t s)
{
    const int g = 1024;
    hipLaunchKernelGGL(k_factor_init, dim3(g), dim3(256), 0, s, a.n, a.dep, a.head, a.nfill, a.queue, a.level, a.out_cnt);
    hipLaunchKernelGGL(k_factor_deps, dim3(g), dim3(256), 0, s, a.n, a.lo_ptr, a.lo_row, a.dep);
    hipLaunchKernelGGL(k_factor_seed_queue, dim3(g), dim3(256), 0, s, a.n, a.dep, a.queue, a.ctr);
    hipLaunchKernelGGL(k_factor, dim3(num_blocks), dim3(kWave), 0, s, a);
}

int factor_lds_capacity() { return kLdsCap; }
int factor_scratch_bytes_per_entry() { return kBytesPerEntry; }

int factor_occupancy_blocks(int device)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_factor, kWave, 0) != hipSuccess) per_cu = 8;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) cus = 256;
    if (per_cu < 1) per_cu = 1;
    return per_cu * cus;
}

}  // namespace rc
