// factor.hip -- randomized elimination on gfx950, one round per DAG level.
//
// Semantics follow cpu_implementation/driver_local.cpp:302-707 and
// gpu_implementation/driver.cu:43-1302 (the reference's per-column step):
// gather the column's strict-lower entries and pending fill edges, merge
// duplicate rows (summing weight and multiplicity), sort by weight, take the
// forward cumulative sums C and S = C[d-1], and for every entry i < d-1 draw
// u in (0,1], pick the first j > i with C[j] - C[i] >= u (S - C[i]) and emit
// the fill edge (max, min, w_i (S - C[i]) / S) into column min; the factor
// column holds -w/S and the pivot is S.  The last column is never eliminated
// (driver.cu:184-197): it keeps count 0 and pivot 0.
//
// MI355X design:
//   * rounds: one launch eliminates every column made ready by the previous
//     launch (one wave64 per column, grid-stride over the round's list), so
//     the kernel boundary orders all communication -- no fences, no waits;
//   * gather, duplicate merge, sorts and the prefix sum run in LDS (global
//     per-wave scratch for columns above the LDS capacity);
//   * the sampling draw is Philox4x32-10 keyed by (seed, i, column): with the
//     canonical (row, weight) merge order and (weight, row) sampling order the
//     factor is a pure function of (matrix, seed), independent of schedule;
//   * each neighbour's dependency counter gets ONE atomic add of
//     (new fills with it as the larger end) - (retired multiplicity); the
//     column whose add reaches zero appends the neighbour to the next round's
//     list with one wave-aggregated atomic;
//   * a column's pending fill records hang off kFillLists linked lists
//     (record index mod kFillLists), so its gather walks them in lockstep,
//     one lane per list, instead of chasing one long chain;
//   * fill records and factor entries are carved from per-wave chunks of two
//     global pools, so pool counters see one atomic per few thousand entries;
//   * pool or scratch exhaustion raises a status bit, later rounds drain, and
//     the host regrows and reruns (same result).
#include "common.cuh"
#include "kernels.h"

namespace rc {

namespace {

constexpr int kLdsCap = 128;                                   // entries per wave in LDS (round kernel)
static_assert(kLdsCap <= 2 * 64, "the round kernel sorts its columns in registers (two slots per lane)");
constexpr int kBytesPerEntry = 16 + 8 + 8 + 4 + 4 + 4 + 4;  // s mw C mrow mmult minc jt
constexpr long long kFillChunk = 2048;
constexpr long long kOutChunk = 2048;

// An entry being sorted: row, payload (multiplicity, later the merged index)
// and weight in one 16-byte record, so a comparator moves two records with
// one LDS instruction each instead of six.
struct __attribute__((aligned(16))) SEnt {
    int row;
    int pay;
    double w;
};

struct ColBuf {
    SEnt *s;
    double *mw, *C;
    int *mrow, *mmult, *minc, *jt;   // jt: sampled partner of every fill
};

__device__ __forceinline__ ColBuf carve(char *base, int cap)
{
    ColBuf b;
    b.s = reinterpret_cast<SEnt *>(base);
    b.mw = reinterpret_cast<double *>(b.s + cap);
    b.C = b.mw + cap;
    b.mrow = reinterpret_cast<int *>(b.C + cap);
    b.mmult = b.mrow + cap;
    b.minc = b.mmult + cap;
    b.jt = b.minc + cap;
    return b;
}

__device__ __forceinline__ int next_pow2(int v)
{
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Ascending sorting network (bitonic with a mirrored first comparator in
// every merge stage); every comparator puts the minimum at the lower index, so
// positions >= len behave as +infinity and need no padding.
// kByRow: key (row, w); otherwise key (w, row); the payload travels along.
// The comparators of a stage touch disjoint pairs, so each lane reads all of
// its pairs of the stage (kNB = P / 128 of them, at most 8 at a time) before
// it compares and writes any: one LDS round trip per stage instead of one per
// 64 comparators.  kNB is a template argument so no slot is computed idly.
template <bool kByRow>
__device__ __forceinline__ bool key_less(int ra, double wa, int rb, double wb)
{
    return kByRow ? (ra < rb || (ra == rb && wa < wb)) : (wa < wb || (wa == wb && ra < rb));
}

template <bool kByRow, int kNB>
__device__ __forceinline__ void wave_sort_nb(SEnt *e, int len, int P)
{
    const int lane = lane_id();
    const int half = P >> 1;
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t0 = 0; t0 < half; t0 += kNB * kWave) {
                int ia[kNB], la[kNB];
                SEnt ei[kNB], el[kNB];
                bool act[kNB];
#pragma unroll
                for (int b = 0; b < kNB; ++b) {
                    const int t = t0 + b * kWave + lane;
                    const int i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
                    const int l = (j == (k >> 1)) ? (i ^ (k - 1)) : (i + j);
                    act[b] = l < len;                      // t < half always; l > i, so i < len as well
                    ia[b] = i;
                    la[b] = l;
                    // read unconditionally (i, l < P <= the buffer's capacity),
                    // so no branch sits between the reads of the slots
                    ei[b] = e[i];
                    el[b] = e[l];
                }
#pragma unroll
                for (int b = 0; b < kNB; ++b) {
                    if (act[b] && key_less<kByRow>(el[b].row, el[b].w, ei[b].row, ei[b].w)) {
                        e[ia[b]] = el[b];
                        e[la[b]] = ei[b];
                    }
                }
            }
            __syncthreads();
        }
    }
}
// Networks of P >= 256 positions (shorter columns use the register sorts).
template <bool kByRow>
__device__ __forceinline__ void wave_sort(SEnt *e, int len)
{
    const int P = next_pow2(len);
    if (P <= 256) wave_sort_nb<kByRow, 2>(e, len, 256);
    else if (P <= 512) wave_sort_nb<kByRow, 4>(e, len, 512);
    else wave_sort_nb<kByRow, 8>(e, len, P);
}

// The same ordering for len <= 64 R entries held in registers, entry i in
// slot i / 64 of lane i % 64: a bitonic network whose comparator pairs are
// lane pairs (shuffles) for distances below 64 and slot pairs inside a lane
// above, with no LDS round trips between stages.  Positions >= len hold
// +infinity keys.  Ties keep their values (the sorted output is unique: equal
// keys carry equal payloads in both uses).
template <bool kByRow, int R>
__device__ __forceinline__ void lane_sort(SEnt *e, int len)
{
    const int lane = lane_id();
    const int P = next_pow2(len);
    int r[R], y[R];
    double x[R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const int i = s * kWave + lane;
        SEnt v;
        v.row = 0x7fffffff;
        v.pay = 0;
        v.w = __builtin_inf();
        if (i < len) v = e[i];
        r[s] = v.row;
        x[s] = v.w;
        y[s] = v.pay;
    }
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= kWave) {
                const int js = j / kWave;
#pragma unroll
                for (int s = 0; s < R; ++s) {
                    const int t = s ^ js;
                    if (t > s && t < R) {
                        const bool asc = ((s * kWave + lane) & k) == 0;
                        const bool swap = asc ? key_less<kByRow>(r[t], x[t], r[s], x[s])
                                              : key_less<kByRow>(r[s], x[s], r[t], x[t]);
                        if (swap) {
                            const int tr = r[s]; r[s] = r[t]; r[t] = tr;
                            const double tx = x[s]; x[s] = x[t]; x[t] = tx;
                            const int ty = y[s]; y[s] = y[t]; y[t] = ty;
                        }
                    }
                }
            } else {
#pragma unroll
                for (int s = 0; s < R; ++s) {
                    const int ro = __shfl_xor(r[s], j, kWave);
                    const double xo = __shfl_xor(x[s], j, kWave);
                    const int yo = __shfl_xor(y[s], j, kWave);
                    const bool asc = ((s * kWave + lane) & k) == 0;
                    const bool keep_min = asc == ((lane & j) == 0);
                    const bool take = keep_min ? key_less<kByRow>(ro, xo, r[s], x[s])
                                               : key_less<kByRow>(r[s], x[s], ro, xo);
                    if (take) { r[s] = ro; x[s] = xo; y[s] = yo; }
                }
            }
        }
    }
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const int i = s * kWave + lane;
        if (i < len) {
            SEnt v;
            v.row = r[s];
            v.pay = y[s];
            v.w = x[s];
            e[i] = v;
        }
    }
    __syncthreads();
}

// kMaxR bounds the register slots (and so the kernel's register budget).  Both
// kernels use 2: above 128 entries the LDS network measured faster than
// 4 or 8 slots per lane.
// kShort (the round kernel): the caller guarantees len <= kMaxR * 64 (longer
// columns are deferred), so only the register networks are compiled in (the
// LDS network's registers would lower the round kernel's occupancy).
// Otherwise (the large-LDS kernels) only the LDS network is.
template <bool kByRow, int kMaxR, bool kShort = false>
__device__ __forceinline__ void column_sort(SEnt *e, int len)
{
    if constexpr (kShort) {
        if (len <= kWave) lane_sort<kByRow, 1>(e, len);
        else lane_sort<kByRow, kMaxR>(e, len);
        return;
    }
    // the large-LDS kernels order columns of up to kRankMax entries by
    // counting and come here for them only when a row run is longer than the
    // in-register merge takes, so they need no register network (its code
    // would only enlarge the kernels' instruction footprint)
    wave_sort<kByRow>(e, len);
}

// ---- counting sorts for the kernels that run one wave per CU
// A column of at most kRankMax entries is put in order by counting, for every
// entry, the keys below its own: each lane compares its kRankSlots-or-fewer
// keys with every key of the column (broadcast LDS reads, kRankBatch keys in
// flight at a time) and then stores each entry at its count: one LDS round
// trip per kRankBatch keys, where the network needs one per stage (21-36
// stages).  The large-LDS kernels run one wave per CU, so the VALU is
// theirs; the round kernel keeps its register networks (its six waves per
// SIMD share the VALU).  Above two keys per lane the comparisons cost more
// than the network's round trips (measured: 129-256 entries 11.3 -> 13.5 us
// for the row order, 9.7 -> 12.9 us for the value order).
constexpr int kRankMax = 128;
constexpr int kRankSlots = kRankMax / kWave;
static_assert(kRankSlots == 2, "the dispatchers take one or two keys per lane");
constexpr int kRun = 8;    // duplicates of one row merged in registers; a longer run uses the network

// ascending 8-input sorting network (19 comparators in 6 layers)
__device__ __forceinline__ void sort8(double *v)
{
    auto cx = [&](int x, int y) {
        const double lo = fmin(v[x], v[y]);
        const double hi = fmax(v[x], v[y]);
        v[x] = lo;
        v[y] = hi;
    };
    cx(0, 2); cx(1, 3); cx(4, 6); cx(5, 7);
    cx(0, 4); cx(1, 5); cx(2, 6); cx(3, 7);
    cx(0, 1); cx(2, 3); cx(4, 5); cx(6, 7);
    cx(2, 4); cx(3, 5);
    cx(1, 4); cx(3, 6);
    cx(1, 2); cx(3, 4); cx(5, 6);
}

// Entries [0, m) of b.s reordered by (row, gathered index): the
// key (row << 32 | index) is unique, so the counts are the positions.  The
// keys are staged in b.C (unused until the scan), padded with +infinity to a
// multiple of kRankBatch; every batch of keys is read before any is compared,
// so a batch costs one LDS round trip.
constexpr int kRankBatch = 16;
__device__ __forceinline__ int rank_pad(int m) { return (m + kRankBatch - 1) & ~(kRankBatch - 1); }

template <int R>
__device__ __forceinline__ void rank_by_row(const ColBuf &b, int m)
{
    const int lane = lane_id();
    unsigned long long *key = reinterpret_cast<unsigned long long *>(b.C);
    int r[R], mu[R], rk[R];
    double w[R];
    unsigned long long k[R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const int i = s * kWave + lane;
        const bool in = i < m;
        SEnt v;
        v.row = 0;
        v.pay = 0;
        v.w = 0.0;
        if (in) v = b.s[i];
        r[s] = v.row;
        w[s] = v.w;
        mu[s] = v.pay;
        k[s] = in ? (((unsigned long long)(unsigned int)r[s] << 32) | (unsigned int)i) : ~0ull;
        if (in) key[i] = k[s];
        rk[s] = 0;
    }
    const int mp = rank_pad(m);
    if (m + lane < mp) key[m + lane] = ~0ull;
    __syncthreads();
    const ulonglong2 *kp = reinterpret_cast<const ulonglong2 *>(key);
    for (int o = 0; o < mp; o += kRankBatch) {
        ulonglong2 kk[kRankBatch / 2];
#pragma unroll
        for (int q = 0; q < kRankBatch / 2; ++q) kk[q] = kp[o / 2 + q];
#pragma unroll
        for (int q = 0; q < kRankBatch / 2; ++q) {
#pragma unroll
            for (int s = 0; s < R; ++s) rk[s] += (kk[q].x < k[s] ? 1 : 0) + (kk[q].y < k[s] ? 1 : 0);
        }
    }
    __syncthreads();   // every entry is in registers: the arrays are rewritten in place
#pragma unroll
    for (int s = 0; s < R; ++s) {
        if (s * kWave + lane < m) {
            SEnt v;
            v.row = r[s];
            v.pay = mu[s];
            v.w = w[s];
            b.s[rk[s]] = v;
        }
    }
    __syncthreads();
}

// Step 2 by counting: entries ordered by (row, index), then every run of one
// row (at most kRun entries) is put in ascending weight in registers and
// summed left to right -- the sums of the (row, weight) sort and merge, bit
// for bit (equal weights are interchangeable).  Returns the merged count, or
// -1 if some run is longer than kRun (the caller then sorts by network).
__device__ __forceinline__ int rank_merge(const ColBuf &b, int m_in)
{
    const int lane = lane_id();
    const int m = __builtin_amdgcn_readfirstlane(m_in);
    const int R = (m + kWave - 1) / kWave;
    if (R <= 1) rank_by_row<1>(b, m);
    else rank_by_row<kRankSlots>(b, m);
    int d = 0;
    bool over = false;
    for (int base_i = 0; base_i < m; base_i += kWave) {
        const int i = base_i + lane;
        bool head = false;
        int rr = 0, ms = 0;
        double ws = 0.0;
        if (i < m) {
            rr = b.s[i].row;
            head = (i == 0) || (b.s[i - 1].row != rr);
        }
        if (head) {
            // the run is [i, i + g): rows are in order; reads past m stay
            // inside the buffer (m + kRun <= its capacity) and are masked
            double v[kRun];
            int g = 0;
#pragma unroll
            for (int t = 0; t < kRun; ++t) {
                const int e = i + t;
                const SEnt x = b.s[e];
                const bool in = e < m && x.row == rr;
                v[t] = in ? x.w : __builtin_inf();
                ms += in ? x.pay : 0;
                g += in ? 1 : 0;
            }
            over = over || (i + kRun < m && b.s[i + kRun].row == rr);
            sort8(v);
            ws = v[0];
#pragma unroll
            for (int t = 1; t < kRun; ++t)
                if (t < g) ws += v[t];
        }
        const uint64_t mask = ballot(head);
        if (head) {
            const int pos = d + lanes_below(mask);
            b.mrow[pos] = rr;
            b.mw[pos] = ws;
            b.mmult[pos] = ms;
            b.minc[pos] = 0;
        }
        d += __popcll(mask);
    }
    const bool any_over = ballot(over) != 0ull;
    __syncthreads();
    return any_over ? -1 : d;
}

// Step 3's order by counting: merged entries [0, d) by (w, row) (rows are
// distinct, so the order is strict), written to b.s (payload = merged
// index).  mw / mrow must be padded with +infinity / INT_MAX up to
// rank_pad(d).
template <int R>
__device__ __forceinline__ void rank_by_value(const ColBuf &b, int d)
{
    const int lane = lane_id();
    int r[R], rk[R];
    double w[R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const int i = s * kWave + lane;
        r[s] = i < d ? b.mrow[i] : 0x7fffffff;
        w[s] = i < d ? b.mw[i] : __builtin_inf();
        rk[s] = 0;
    }
    const int dp = rank_pad(d);
    const double2 *wp = reinterpret_cast<const double2 *>(b.mw);
    const int2 *rp = reinterpret_cast<const int2 *>(b.mrow);
    for (int o = 0; o < dp; o += kRankBatch) {
        double2 wo[kRankBatch / 2];
        int2 ro[kRankBatch / 2];
#pragma unroll
        for (int q = 0; q < kRankBatch / 2; ++q) {
            wo[q] = wp[o / 2 + q];
            ro[q] = rp[o / 2 + q];
        }
#pragma unroll
        for (int q = 0; q < kRankBatch / 2; ++q) {
#pragma unroll
            for (int s = 0; s < R; ++s) {
                rk[s] += (wo[q].x < w[s] || (wo[q].x == w[s] && ro[q].x < r[s])) ? 1 : 0;
                rk[s] += (wo[q].y < w[s] || (wo[q].y == w[s] && ro[q].y < r[s])) ? 1 : 0;
            }
        }
    }
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const int i = s * kWave + lane;
        if (i < d) {
            SEnt v;
            v.row = r[s];
            v.pay = i;
            v.w = w[s];
            b.s[rk[s]] = v;
        }
    }
    __syncthreads();
}

// Step 2 by the sorting networks: (row, w) order, then each run of one row
// summed left to right (ascending weight).  Returns the merged count.
template <int kMaxR, bool kShort>
__device__ __forceinline__ int sort_merge(const ColBuf &b, int m)
{
    const int lane = lane_id();
    column_sort<true, kMaxR, kShort>(b.s, m);
    int d = 0;
    for (int base_i = 0; base_i < m; base_i += kWave) {
        int i = base_i + lane;
        bool head = false;
        int r = 0, ms = 0;
        double ws = 0.0;
        if (i < m) {
            r = b.s[i].row;
            head = (i == 0) || (b.s[i - 1].row != r);
            if (head) {
                // the run [i, i + g) summed left to right, kRun records per
                // LDS round trip (records past m are masked)
                ws = b.s[i].w;
                ms = b.s[i].pay;
                bool more = true;
                for (int t0 = i + 1; more; t0 += kRun) {
                    SEnt x[kRun];
#pragma unroll
                    for (int u = 0; u < kRun; ++u) x[u] = b.s[t0 + u < m ? t0 + u : m - 1];
#pragma unroll
                    for (int u = 0; u < kRun; ++u) {
                        more = more && t0 + u < m && x[u].row == r;
                        if (more) {
                            ws += x[u].w;
                            ms += x[u].pay;
                        }
                    }
                }
            }
        }
        uint64_t mask = ballot(head);
        if (head) {
            int pos = d + lanes_below(mask);
            b.mrow[pos] = r;
            b.mw[pos] = ws;
            b.mmult[pos] = ms;
            b.minc[pos] = 0;
        }
        d += __popcll(mask);
    }
    __syncthreads();
    return d;
}

__device__ __forceinline__ void raise_status(unsigned long long *ctr, unsigned long long bits)
{
    atomicOr(&ctr[kCtrStatus], bits);
}

struct Cursor {
    long long cur, end;
};

// Wave-uniform carve of `count` entries from a per-wave chunk of a pool.
__device__ __forceinline__ bool carve_pool(Cursor &c, long long count, long long chunk, unsigned long long *top,
                                           long long cap, long long &base)
{
    if (c.cur + count > c.end) {
        long long size = count > chunk ? count : chunk;
        long long b = 0;
        if (lane_id() == 0) b = (long long)atomicAdd(top, (unsigned long long)size);
        b = bcast(b, 0);
        if (b + size > cap) return false;
        c.cur = b;
        c.end = b + size;
    }
    base = c.cur;
    c.cur += count;
    return true;
}

// ---- segmented ready lists ------------------------------------------------
__device__ __forceinline__ unsigned int *list_set(const FactorArgs &a, int round)
{
    return a.seg_count + (size_t)(round % 3) * kListSlots;
}

// Lane q: effective entry count c of slot q of a list and the inclusive
// prefix over slots; the list's length is the prefix at lane 63.
__device__ __forceinline__ void list_counts(const FactorArgs &a, const unsigned int *set, unsigned int &c,
                                            unsigned int &incl)
{
    const int lane = lane_id();
    const unsigned int v = set[lane];
    c = lane < kListSegs ? min(v, (unsigned int)a.seg_cap) : v;
    incl = c;
    for (int off = 1; off < kWave; off <<= 1) {
        const unsigned int o = __shfl_up(incl, off, kWave);
        if (lane >= off) incl += o;
    }
}

// Storage index of entry i (wave-uniform, below the length) of a list.
__device__ __forceinline__ size_t list_index(const FactorArgs &a, unsigned int c, unsigned int incl, unsigned int i)
{
    const int s = __popcll(ballot(incl <= i));          // slots wholly before entry i
    const unsigned int excl = __shfl(incl - c, s, kWave);
    return (size_t)s * (size_t)a.seg_cap + (i - excl);
}

// Appends `value` of every lane with `ready` to the list of counter set `set`:
// one atomic on this workgroup's segment counter, and one on the overflow
// counter for entries past the segment's capacity.
__device__ __forceinline__ void list_append(const FactorArgs &a, unsigned int *set, int *list, bool ready, int value)
{
    const uint64_t mask = ballot(ready);
    if (!mask) return;
    const unsigned int cnt = (unsigned int)__popcll(mask);
    const unsigned int cap = (unsigned int)a.seg_cap;
    const int seg = (int)(blockIdx.x % kListSegs);
    unsigned int base = 0;
    if (lane_id() == 0) base = atomicAdd(&set[seg], cnt);
    base = bcast(base, 0);
    const unsigned int pos = base + (unsigned int)lanes_below(mask);
    if (base + cnt > cap) {
        const unsigned int lo = base > cap ? base : cap;
        unsigned int ob = 0;
        if (lane_id() == 0) ob = atomicAdd(&set[kListSegs], base + cnt - lo);
        ob = bcast(ob, 0);
        if (ready && pos >= cap) list[(size_t)kListSegs * cap + ob + (pos - lo)] = value;
    }
    if (ready && pos < cap) list[(size_t)seg * cap + pos] = value;
}

// Gathers column k's strict-lower entries and its pending fills (kFillLists
// linked lists, walked in lockstep, one lane per list; f = this lane's list
// head) into b, writing the first `cap` entries; returns the total count, or
// with kStopOverCap any count above `cap` as soon as one is reached (the
// column is then handed on without walking the rest of its lists).
// The order of the gathered entries does not matter: the (row, weight) sort
// makes it canonical.
template <bool kStopOverCap>
__device__ __forceinline__ int gather_column(const FactorArgs &a, int lo0, int norig, int f, const ColBuf &b, int cap)
{
    const int lane = lane_id();
    for (int t = lane; t < norig && t < cap; t += kWave) {
        SEnt v;
        v.row = a.lo_row[lo0 + t];
        v.pay = 1;
        v.w = a.lo_w[lo0 + t];
        b.s[t] = v;
    }
    int t = norig;
    while (true) {
        if (kStopOverCap && t > cap) break;
        const bool act = f >= 0;
        const uint64_t mask = ballot(act);
        if (!mask) break;
        if (act) {
            const FillRec rec = a.pool[f];
            const int pos = t + lanes_below(mask);
            if (pos < cap) {
                SEnt v;
                v.row = rec.row;
                v.pay = 1;
                v.w = rec.w;
                b.s[pos] = v;
            }
            f = rec.next;
        }
        t += __popcll(mask);
    }
    __syncthreads();
    return t;
}

// Steps 2-5 of a column whose m >= 1 entries are gathered in b.  Inlined at
// every call so that b's accesses compile to LDS or global instructions (a
// generic pointer would make each LDS access wait for all memory traffic).
template <int kMaxR, bool kAsync, bool kShort = false>
__device__ __forceinline__ bool process_column(const FactorArgs &a, int k, int round, const ColBuf &b, int m,
                                               Cursor &fcur, Cursor &ocur, int *next, unsigned int *next_set)
{
    const int lane = lane_id();
    const int last = a.n - 1;
    // ---- 2. sort by (row, w); merge duplicates left to right (:520-560)
    int d = -1;
    if constexpr (!kShort) {
        if (m <= kRankMax) d = rank_merge(b, m);
    }
    if (d < 0) d = sort_merge<kMaxR, kShort>(b, m);

    // ---- 3. sort merged entries by (w, row), carrying the merged index;
    //         forward cumulative sums (:573-593)
    bool ranked = false;
    if constexpr (!kShort) {
        if (d <= kRankMax) {
            d = __builtin_amdgcn_readfirstlane(d);
            if (d + lane < rank_pad(d)) {
                b.mw[d + lane] = __builtin_inf();
                b.mrow[d + lane] = 0x7fffffff;
            }
            __syncthreads();
            const int R = (d + kWave - 1) / kWave;
            if (R <= 1) rank_by_value<1>(b, d);
            else rank_by_value<kRankSlots>(b, d);
            ranked = true;
        }
    }
    if (!ranked) {
        for (int t = lane; t < d; t += kWave) {
            SEnt v;
            v.row = b.mrow[t];
            v.pay = t;
            v.w = b.mw[t];
            b.s[t] = v;
        }
        __syncthreads();
        column_sort<false, kMaxR, kShort>(b.s, d);
    }
    // canonical blocked scan (shared with the oracle): per 64-entry block the
    // doubling scan x_i += x_{i-off}, off = 1..32, then the previous block's
    // last sum is added
    {
        double carry = 0.0;
        for (int base_i = 0; base_i < d; base_i += kWave) {
            const int i = base_i + lane;
            double x = i < d ? b.s[i].w : 0.0;
            for (int off = 1; off < kWave; off <<= 1) {
                const double o = __shfl_up(x, off, kWave);
                if (lane >= off) x = x + o;
            }
            const double c = base_i == 0 ? x : carry + x;
            if (i < d) b.C[i] = c;
            carry = __shfl(c, kWave - 1, kWave);
        }
    }
    __syncthreads();
    const double S = b.C[d - 1];

    // ---- 4. sampled partners (:598-631; gpu auxilliary.hpp:543-592):
    //         partner j of every fill i and the new dependencies per neighbour
    const int nfill_out = d - 1;
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
        b.jt[i] = lo;
        const SEnt ei = b.s[i], el = b.s[lo];
        atomicAdd(&b.minc[ei.row > el.row ? ei.pay : el.pay], 1);   // the larger row gains a dependency
    }
    long long fbase = 0, obase = 0;
    if (nfill_out > 0 && !carve_pool(fcur, nfill_out, kFillChunk, &a.ctr[kCtrFillTop], a.pool_cap, fbase)) {
        if (lane == 0) raise_status(a.ctr, kStatusFillPool);
        return false;
    }
    if (!carve_pool(ocur, d, kOutChunk, &a.ctr[kCtrOutTop], a.out_cap, obase)) {
        if (lane == 0) raise_status(a.ctr, kStatusOutPool);
        return false;
    }
    // a column is eliminated in the round after its last dependency, so its
    // level in the triangular-solve DAG is its round (:638-645)
    if (lane == 0) {
        if (!kAsync) a.level[k] = round;     // the queue-driven tail's levels come from atomicMax
        a.out_start[k] = obase;
        a.out_cnt[k] = d;
        a.diag[k] = S;
    }
    __syncthreads();   // minc and jt complete

    // ---- 5. per block of 64 entries: the neighbour's net dependency update
    //         (:653-687), the fill edge (max, min, w_i (S - C_i) / S) pushed
    //         onto column min's lists, and the factor entry -w/S in row order.
    //         Up to kB blocks are handled together: their returning atomics
    //         (list-head exchanges, dependency adds) are all in flight at once,
    //         so a long column pays one memory round trip per kB blocks, not
    //         one per block.  The queue-driven tail makes every push, entry and
    //         raised level visible (release fence) before any dependency add,
    //         since a neighbour may be taken by another wave as soon as its
    //         count reaches zero.
    constexpr int kB = kShort ? 2 : 8;
    for (int sb = 0; sb < d; sb += kB * kWave) {
        // the returning atomics of the batch: list-head exchanges and (rounds
        // only: nothing orders them there but the kernel boundary) the
        // dependency adds, all in flight together
        int old[kB], r[kB], nd[kB];
        bool upd[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const int e = sb + u * kWave + lane;
            old[u] = -1;
            r[u] = -1;
            upd[u] = false;
            nd[u] = 1;
            if (e < nfill_out) {
                const int j = b.jt[e];
                const int ri = b.s[e].row, rj = b.s[j].row;
                const int cc = ri > rj ? rj : ri;
                const int rec = (int)(fbase + e);
                old[u] = atomicExch(&a.head[(size_t)cc * kFillLists + (rec & (kFillLists - 1))], rec);
            }
            if (!kAsync && e < d) {
                r[u] = b.mrow[e];
                const int delta = b.minc[e] - b.mmult[e];
                upd[u] = r[u] != last && delta != 0;
                if (upd[u]) nd[u] = atomicAdd(&a.dep[r[u]], delta) + delta;
            }
        }
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const int e = sb + u * kWave + lane;
            if (e < d) {
                const int rr = b.mrow[e];
                a.out_row[obase + e] = rr;
                a.out_val[obase + e] = -(b.mw[e] / S);
                if (kAsync && rr != last) atomicMax(&a.level[rr], round + 1);
            }
            if (e < nfill_out) {
                const int j = b.jt[e];
                const SEnt se = b.s[e];
                const int ri = se.row, rj = b.s[j].row;
                FillRec fr;
                fr.row = ri > rj ? ri : rj;
                fr.next = old[u];
                fr.w = (se.w * (S - b.C[e])) / S;
                a.pool[(int)(fbase + e)] = fr;
            }
        }
        if constexpr (!kAsync) {
#pragma unroll
            for (int u = 0; u < kB; ++u) list_append(a, next_set, next, upd[u] && nd[u] == 0, r[u]);
        }
    }
    if constexpr (kAsync) {
        fence_release();
        for (int sb = 0; sb < d; sb += kB * kWave) {
            int r[kB], nd[kB];
            bool upd[kB];
#pragma unroll
            for (int u = 0; u < kB; ++u) {
                const int e = sb + u * kWave + lane;
                r[u] = -1;
                upd[u] = false;
                nd[u] = 1;
                if (e < d) {
                    r[u] = b.mrow[e];
                    const int delta = b.minc[e] - b.mmult[e];
                    upd[u] = r[u] != last && delta != 0;
                    if (upd[u]) nd[u] = atomicAdd(&a.dep[r[u]], delta) + delta;
                }
            }
            fence_acq_rel();
            // the ready neighbours of these blocks, one queue atomic for all
            unsigned int cnt = 0;
            uint64_t mk[kB];
#pragma unroll
            for (int u = 0; u < kB; ++u) {
                mk[u] = ballot(upd[u] && nd[u] == 0);
                cnt += (unsigned int)__popcll(mk[u]);
            }
            if (cnt) {
                unsigned int qb = 0;
                if (lane == 0) qb = atomicAdd(a.aq_ctr + kAqTail, cnt);
                qb = bcast(qb, 0);
#pragma unroll
                for (int u = 0; u < kB; ++u) {
                    // every remaining column is queued once: a slot past the
                    // queue's end means inconsistent dependency counts
                    const bool ready = (mk[u] >> lane) & 1ull;
                    const unsigned int slot = qb + (unsigned int)lanes_below(mk[u]);
                    if (ready && slot < a.aq_total) st_relaxed(&next[slot], r[u]);
                    if (ready && slot >= a.aq_total) raise_status(a.ctr, kStatusInconsistent);
                    qb += (unsigned int)__popcll(mk[u]);
                }
            }
        }
    }
    __syncthreads();
    return true;
}

// Eliminates column k; newly ready neighbours are appended to `next`.
// Returns false if the pools or the scratch are exhausted (status set).
// kDefer (round kernel): a column larger than the LDS buffer is appended to
// the round's deferred list; otherwise (large-LDS kernel) it is gathered
// again into this workgroup's global scratch.
template <int kMaxR, bool kDefer, bool kAsync = false>
__device__ bool eliminate_column(const FactorArgs &a, int k, int round, char *lds, int lds_cap, Cursor &fcur,
                                 Cursor &ocur, int *next, unsigned int *next_set, unsigned int *defer_count)
{
    const int lane = lane_id();
    // the list heads do not depend on the column's slice: one memory trip for both
    const int f0 = lane < kFillLists ? a.head[(size_t)k * kFillLists + lane] : -1;
    const int lo0 = a.lo_ptr[k];
    const int norig = a.lo_ptr[k + 1] - lo0;

    // ---- 1. gather (driver_local.cpp:459-487; cpu auxilliary.hpp:89-173)
    // into this kernel's LDS; the entries are counted as they come (no fill
    // counter is kept), and a column that turns out larger than the buffer is
    // deferred (round kernel) or gathered again into the global scratch
    const ColBuf b = carve(lds, lds_cap);
    int m = gather_column<kDefer>(a, lo0, norig, f0, b, lds_cap);
    if (m == 0) {                      // empty column: zero pivot (driver_local.cpp:491-498)
        if (lane == 0) {
            if (!kAsync) a.level[k] = round;
            a.out_start[k] = 0;
            a.out_cnt[k] = 0;
            a.diag[k] = 0.0;
        }
        return true;
    }
    const int P = next_pow2(m);
    if (P > lds_cap) {
        if constexpr (kDefer) {
            // too big for this kernel's LDS: the round's large-LDS kernel takes it
            if (lane == 0) a.big_list[atomicAdd(defer_count, 1u)] = k;
            return true;
        } else {
            if (P > a.scratch_cap) {
                if (lane == 0) {
                    atomicMax(&a.ctr[kCtrScratchNeed], (unsigned long long)P);
                    raise_status(a.ctr, kStatusScratch);
                }
                return false;
            }
            const ColBuf g = carve(a.scratch + (size_t)blockIdx.x * (size_t)a.scratch_cap * kBytesPerEntry, a.scratch_cap);
            if (gather_column<false>(a, lo0, norig, f0, g, a.scratch_cap) != m) {
                if (lane == 0) raise_status(a.ctr, kStatusInconsistent);
                return false;
            }
            return process_column<kMaxR, kAsync>(a, k, round, g, m, fcur, ocur, next, next_set);
        }
    }
    // the round kernel (kDefer) reaches here only with P <= lds_cap = kLdsCap entries
    return process_column<kMaxR, kAsync, kDefer>(a, k, round, b, m, fcur, ocur, next, next_set);
}

}  // namespace

// One elimination round: every column of `cur` (count counts[round]).
__global__ void __launch_bounds__(kWave) k_factor_round(const FactorArgs *__restrict__ pa, int round, const int *cur,
                                                        int *next)
{
    const FactorArgs &a = *pa;
    __shared__ __attribute__((aligned(16))) char lds[kLdsCap * kBytesPerEntry];
    // this round's list counters (one per lane) with the status in one trip;
    // workgroup 0 clears the counter set of round + 2 (last used by round - 1)
    // and records the round's length for the host
    unsigned int c, incl;
    list_counts(a, list_set(a, round), c, incl);
    const unsigned long long status = a.ctr[kCtrStatus];
    const unsigned int ncur = __shfl(incl, kWave - 1, kWave);
    if (blockIdx.x == 0) {
        list_set(a, round + 2)[lane_id()] = 0u;
        if (lane_id() == 0) a.round_count[round] = ncur;
    }
    if (ncur <= (unsigned int)kBigGrid) return;     // a narrow round: the large-LDS kernel takes all of it
    if (status != 0ull) return;
    Cursor fcur, ocur;
    long long *ws = a.wave_state + 4 * (size_t)blockIdx.x;
    fcur.cur = ws[0]; fcur.end = ws[1];
    ocur.cur = ws[2]; ocur.end = ws[3];
    unsigned int *next_set = list_set(a, round + 1);
    bool ok = true;
    for (unsigned int i = blockIdx.x; i < ncur; i += gridDim.x) {
        const int col = cur[list_index(a, c, incl, i)];
        if (!eliminate_column<2, true>(a, col, round, lds, kLdsCap, fcur, ocur, next, next_set,
                                       &a.big_count[round])) { ok = false; break; }
    }
    (void)ok;
    if (lane_id() == 0) {
        ws[0] = fcur.cur; ws[1] = fcur.end;
        ws[2] = ocur.cur; ws[3] = ocur.end;
    }
}

// The round's columns deferred by k_factor_round (more than kLdsCap entries):
// one wave per column with kBigCap entries of LDS (columns above that use the
// per-workgroup global scratch).  Runs after k_factor_round of the same round.
__global__ void __launch_bounds__(kWave) k_factor_round_big(const FactorArgs *__restrict__ pa, int round, const int *cur,
                                                            int *next)
{
    const FactorArgs &a = *pa;
    extern __shared__ __attribute__((aligned(16))) char lds_big[];
    // a round of at most kBigGrid columns is eliminated here entirely (one
    // column per workgroup); otherwise the columns k_factor_round deferred
    unsigned int c, incl;
    list_counts(a, list_set(a, round), c, incl);
    const unsigned int ndef = a.big_count[round];
    const unsigned long long status = a.ctr[kCtrStatus];
    const unsigned int ncur = __shfl(incl, kWave - 1, kWave);
    const bool narrow = ncur <= (unsigned int)kBigGrid;
    const unsigned int nbig = narrow ? ncur : ndef;
    if (nbig == 0) return;
    if (status != 0ull) return;
    Cursor fcur, ocur;
    long long *ws = a.wave_state + 4 * ((size_t)a.big_ws + blockIdx.x);
    fcur.cur = ws[0]; fcur.end = ws[1];
    ocur.cur = ws[2]; ocur.end = ws[3];
    for (unsigned int i = blockIdx.x; i < nbig; i += gridDim.x) {
        const int col = narrow ? cur[list_index(a, c, incl, i)] : a.big_list[i];
        if (!eliminate_column<2, false>(a, col, round, lds_big, kBigCap, fcur, ocur, next, list_set(a, round + 1), nullptr))
            break;
    }
    if (lane_id() == 0) {
        ws[0] = fcur.cur; ws[1] = fcur.end;
        ws[2] = ocur.cur; ws[3] = ocur.end;
    }
}

// ---- queue-driven tail -----------------------------------------------------
// Once a round holds at most kBigGrid columns, the rest of the elimination runs
// in this one persistent launch: every wave takes the next queue index, waits
// until that slot holds a column (slots start at -1), eliminates it in the
// large LDS buffer and appends the neighbours it made ready.  A column's level
// is 1 + the largest level of its lower neighbours (atomicMax by each of them
// before its dependency update), the same value the rounds assign.  Slot h is
// written by the wave that made its column ready, which is running, so every
// wait ends; the waits are bounded all the same (status bit on expiry).
__device__ __forceinline__ bool factor_should_stop(const FactorArgs &a)
{
    return ld_relaxed(&a.ctr[kCtrStatus]) != 0ull;
}

__global__ void __launch_bounds__(kWave) k_factor_async(const FactorArgs *__restrict__ pa, unsigned int total)
{
    const FactorArgs &a = *pa;
    extern __shared__ __attribute__((aligned(16))) char lds_big[];
    const int lane = lane_id();
    Cursor fcur, ocur;
    long long *ws = a.wave_state + 4 * ((size_t)a.big_ws + blockIdx.x);
    fcur.cur = ws[0]; fcur.end = ws[1];
    ocur.cur = ws[2]; ocur.end = ws[3];
    while (true) {
        unsigned int h = 0;
        if (lane == 0) h = atomicAdd(a.aq_ctr + kAqHead, 1u);
        h = bcast(h, 0);
        if (h >= total) break;
        int k = -1;
        if (lane == 0) {
            k = ld_relaxed(&a.aq[h]);
            if (k < 0) {
                const long long t0 = wall_clock64();
                int nap = 1;
                while ((k = ld_relaxed(&a.aq[h])) < 0) {
                    if (factor_should_stop(a)) break;
                    if (wall_clock64() - t0 > a.spin_ticks) {
                        raise_status(a.ctr, kStatusTimeout);
                        break;
                    }
                    for (int t = 0; t < nap; ++t) __builtin_amdgcn_s_sleep(2);
                    nap = nap < 8 ? nap * 2 : 8;
                }
            }
        }
        k = bcast(k, 0);
        if (k < 0) break;                          // stopped (status raised)
        fence_acquire();
        const int lvl = a.level[k];
        if (!eliminate_column<2, false, true>(a, k, lvl, lds_big, kBigCap, fcur, ocur, a.aq, nullptr, nullptr))
            break;
        if (factor_should_stop(a)) break;
    }
    if (lane == 0) {
        ws[0] = fcur.cur; ws[1] = fcur.end;
        ws[2] = ocur.cur; ws[3] = ocur.end;
    }
}

// The queue starts with round `round`'s list (its columns' level is the round).
// One workgroup.  Entry i of the list is found per thread from the slot
// counts kept in LDS (list_index needs a wave-uniform entry).
__global__ void k_factor_async_init(FactorArgs a, int round, const int *cur, unsigned int total)
{
    __shared__ unsigned int s_c[kWave], s_incl[kWave];
    if (threadIdx.x < kWave) {
        unsigned int c, incl;
        list_counts(a, list_set(a, round), c, incl);
        s_c[threadIdx.x] = c;
        s_incl[threadIdx.x] = incl;
    }
    for (unsigned int i = threadIdx.x; i < total; i += blockDim.x) a.aq[i] = -1;
    __syncthreads();
    const unsigned int ncur = min(s_incl[kWave - 1], total);
    for (unsigned int i = threadIdx.x; i < ncur; i += blockDim.x) {
        int s = 0;                                   // slot of entry i: the first with incl > i
        while (s < kWave - 1 && s_incl[s] <= i) ++s;
        const size_t idx = (size_t)s * (size_t)a.seg_cap + (i - (s_incl[s] - s_c[s]));
        const int col = cur[idx];
        a.aq[i] = col;
        a.level[col] = round;
    }
    if (threadIdx.x == 0) {
        a.aq_ctr[kAqHead] = 0u;
        a.aq_ctr[kAqTail] = ncur;
    }
}

// ---- set-up kernels ------------------------------------------------------

__global__ void k_factor_init(int n, const int *dep_init, int *dep, int *head, int *level, int *out_cnt,
                              long long *out_start, double *diag)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < (size_t)n * kFillLists;
         i += (size_t)gridDim.x * blockDim.x)
        head[i] = -1;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        dep[i] = dep_init[i];
        level[i] = 0;
        out_cnt[i] = 0;
        out_start[i] = 0;
        diag[i] = 0.0;
    }
}

// Round 0: every column without lower-index neighbours, except the last.
__global__ void k_factor_seed(FactorArgs a, const int *dep)
{
    for (int base = blockIdx.x * blockDim.x; base < a.n; base += gridDim.x * blockDim.x) {
        const int i = base + threadIdx.x;
        list_append(a, list_set(a, 0), a.list0, i < a.n - 1 && dep[i] == 0, i);
    }
}

void launch_factor_setup(const FactorArgs &a, const int *dep_init, hipStream_t s)
{
    hipLaunchKernelGGL(k_factor_init, dim3(1024), dim3(256), 0, s, a.n, dep_init, a.dep, a.head, a.level,
                       a.out_cnt, a.out_start, a.diag);
    hipLaunchKernelGGL(k_factor_seed, dim3(1024), dim3(256), 0, s, a, (const int *)a.dep);
}

void launch_factor_rounds(const FactorArgs &a, const FactorArgs *da, int num_blocks, int first_round, int nrounds,
                          hipStream_t s)
{
    for (int r = first_round; r < first_round + nrounds; ++r) {
        const int *cur = (r & 1) ? a.list1 : a.list0;
        int *next = (r & 1) ? a.list0 : a.list1;
        hipLaunchKernelGGL(k_factor_round, dim3(num_blocks), dim3(kWave), 0, s, da, r, cur, next);
        hipLaunchKernelGGL(k_factor_round_big, dim3(kBigGrid), dim3(kWave), (size_t)kBigCap * kBytesPerEntry, s, da, r,
                           cur, next);
    }
}

void factor_kernel_setup()
{
    (void)hipFuncSetAttribute((const void *)k_factor_round_big, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)((size_t)kBigCap * kBytesPerEntry));
    (void)hipFuncSetAttribute((const void *)k_factor_async, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)((size_t)kBigCap * kBytesPerEntry));
}

void launch_factor_async(const FactorArgs &a0, const FactorArgs *da, int round, unsigned int total, hipStream_t s)
{
    FactorArgs a = a0;
    a.aq_total = total;            // total <= n - 1 slots of the n-entry queue
    const int *cur = (round & 1) ? a.list1 : a.list0;
    // one block: the init kernel's slot clearing must finish before it seeds the queue
    hipLaunchKernelGGL(k_factor_async_init, dim3(1), dim3(1024), 0, s, a, round, cur, total);
    hipLaunchKernelGGL(k_factor_async, dim3(kBigGrid), dim3(kWave), (size_t)kBigCap * kBytesPerEntry, s, da, total);
}

int factor_lds_capacity() { return kLdsCap; }
int factor_scratch_bytes_per_entry() { return kBytesPerEntry; }
long long factor_pool_slack(int num_blocks) { return (long long)(num_blocks + kBigGrid) * (kFillChunk + kOutChunk); }

int factor_occupancy_blocks(int device)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_factor_round, kWave, 0) != hipSuccess) per_cu = 8;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) cus = 256;
    if (per_cu < 1) per_cu = 1;
    return per_cu * cus;
}

// Statistics after the factor: nnz_off, fills, max degree, zero pivots
// (over the first n-1 columns; the last column is a zero pivot by definition).
__global__ void k_factor_stats(int n, const int *out_cnt, unsigned long long *ctr)
{
    unsigned long long nnz = 0, fills = 0, zp = 0, mx = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        unsigned long long d = (unsigned long long)out_cnt[i];
        nnz += d;
        fills += d > 0 ? d - 1 : 0;
        zp += d == 0 ? 1 : 0;
        mx = d > mx ? d : mx;
    }
    for (int off = 32; off > 0; off >>= 1) {
        nnz += __shfl_xor(nnz, off, kWave);
        fills += __shfl_xor(fills, off, kWave);
        zp += __shfl_xor(zp, off, kWave);
        unsigned long long o = __shfl_xor(mx, off, kWave);
        mx = o > mx ? o : mx;
    }
    if (lane_id() == 0) {
        atomicAdd(&ctr[kCtrNnzOff], nnz);
        atomicAdd(&ctr[kCtrFills], fills);
        atomicAdd(&ctr[kCtrZeroPivots], zp);
        atomicMax(&ctr[kCtrMaxDeg], mx);
    }
}

// Level of the last column (never eliminated): 1 + max level of the columns
// whose factor column holds it (rows are sorted, so it is their last entry).
__global__ void k_last_level(int n, const long long *out_start, const int *out_cnt, const int *out_row,
                             int *level)
{
    int m = -1;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n - 1; k += gridDim.x * blockDim.x) {
        int c = out_cnt[k];
        if (c > 0 && out_row[out_start[k] + c - 1] == n - 1) m = level[k] > m ? level[k] : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        int o = __shfl_xor(m, off, kWave);
        m = o > m ? o : m;
    }
    if (lane_id() == 0 && m >= 0) atomicMax(&level[n - 1], m + 1);
}

void launch_factor_finish(const FactorArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_factor_stats, dim3(1024), dim3(256), 0, s, a.n, (const int *)a.out_cnt, a.ctr);
    if (a.n > 1)
        hipLaunchKernelGGL(k_last_level, dim3(1024), dim3(256), 0, s, a.n, (const long long *)a.out_start,
                           (const int *)a.out_cnt, (const int *)a.out_row, a.level);
}

}  // namespace rc
