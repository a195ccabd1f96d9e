// prep.hip -- the input preparation on the device: symmetric permutation of
// the loaded Laplacian (write_graph.jl:333-365 permutes offline in Julia),
// SDDM ground append (write_graph.jl:367-383), ground trim
// (driver_physics.cu:1595-1734), the strict-lower slice with |values|
// (cpu pre_process.hpp:763-850) and the dependency counts (pre_process.hpp:
// 474-496).  The results equal the host path's (sparse.cpp) entry for entry:
// columns are ordered by (row, then input order), which is the host's
// (row, value) order because the loaded matrix keeps equal rows by value.
#include "common.cuh"
#include "kernels.h"

namespace rc {

namespace {

constexpr int kPB = 256;

// rowk[p] = inv[ind[p]] (the entry's row in the new numbering)
__global__ void k_map_rows(int nnz, const int *ind, const int *inv, uint32_t *rowk)
{
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += gridDim.x * blockDim.x)
        rowk[p] = (uint32_t)inv[ind[p]];
}

// dst[q] = (uint32) src[idx[q]]
__global__ void k_gather_key(const int *src, const int *idx, uint32_t *dst, int n)
{
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
        dst[q] = (uint32_t)src[idx[q]];
}

// out_ind[q] = inv[ind[idx[q]]], out_val[q] = val[idx[q]]
__global__ void k_perm_finish(int nnz, const int *idx, const int *ind, const int *inv, const double *val, int *out_ind,
                              double *out_val)
{
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nnz; q += gridDim.x * blockDim.x) {
        const int p = idx[q];
        out_ind[q] = inv[ind[p]];
        out_val[q] = val[p];
    }
}

// Column sums in row order, one thread per column (bit-equal to the host's
// sequential sum), and the ground entries app = min(0, -sum) with a flag.
__global__ void k_colsum(int n, const int *ptr, const double *val, double *cs, double *app, int *flag)
{
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int p = ptr[j]; p < ptr[j + 1]; ++p) s += val[p];
        cs[j] = s;
        const double v = -s;
        const double a = v > 0.0 ? 0.0 : v;
        app[j] = a;
        flag[j] = a != 0.0 ? 1 : 0;
    }
}

// counts of the ground-appended matrix: column j gains (n, app[j]) when
// flagged; the ground column holds the flagged j and the corner
__global__ void k_ground_count(int n, const int *ptr, const int *flag, int extra, int *cnt)
{
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j <= n; j += gridDim.x * blockDim.x)
        cnt[j] = j < n ? ptr[j + 1] - ptr[j] + flag[j] : extra + 1;
}

// one wave per column: copy the column, then the ground entry
__global__ void k_ground_fill(int n, const int *ptr, const int *ind, const double *val, const double *app,
                              const int *flag, const int *fpos, double corner, const int *optr, int *oind,
                              double *oval)
{
    const int lane = threadIdx.x & 63;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    for (int j = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; j < n; j += nw) {
        const int b = ptr[j], len = ptr[j + 1] - b, o = optr[j];
        for (int t = lane; t < len; t += 64) {
            oind[o + t] = ind[b + t];
            oval[o + t] = val[b + t];
        }
        if (lane == 0 && flag[j]) {
            oind[o + len] = n;
            oval[o + len] = app[j];
            const int g = optr[n] + fpos[j];          // ground column: flagged j in order
            oind[g] = j;
            oval[g] = app[j];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        oind[optr[n + 1] - 1] = n;
        oval[optr[n + 1] - 1] = corner;
    }
}

// trim: column j < n-1 without row n-1
__global__ void k_trim_count(int n, const int *ptr, const int *ind, int *cnt)
{
    const int lane = threadIdx.x & 63;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    for (int j = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; j < n - 1; j += nw) {
        int c = 0;
        for (int p = ptr[j] + lane; p < ptr[j + 1]; p += 64) c += ind[p] != n - 1;
        c = wave_sum_int(c);
        if (lane == 0) cnt[j] = c;
    }
}

__global__ void k_trim_fill(int n, const int *ptr, const int *ind, const double *val, const int *optr, int *oind,
                            double *oval)
{
    const int lane = threadIdx.x & 63;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    for (int j = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; j < n - 1; j += nw) {
        int o = optr[j];
        for (int p0 = ptr[j]; p0 < ptr[j + 1]; p0 += 64) {
            const int p = p0 + lane;
            const bool keep = p < ptr[j + 1] && ind[p] != n - 1;
            const unsigned long long m = __ballot(keep);
            if (keep) {
                const int r = __popcll(m & ((1ull << lane) - 1ull));
                oind[o + r] = ind[p];
                oval[o + r] = val[p];
            }
            o += __popcll(m);
        }
    }
}

// lower slice and dependency counts, one wave per column: rows are sorted,
// so the strict-lower entries of column j are a suffix; dep[j] counts the
// rows above j (the pattern is symmetric: the lower entries in row j)
__global__ void k_lower_count(int n, const int *ptr, const int *ind, int *cnt, int *dep)
{
    const int lane = threadIdx.x & 63;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    for (int j = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; j < n; j += nw) {
        int lo = 0, up = 0;
        for (int p = ptr[j] + lane; p < ptr[j + 1]; p += 64) {
            const int r = ind[p];
            lo += r > j;
            up += r < j;
        }
        lo = wave_sum_int(lo);
        up = wave_sum_int(up);
        if (lane == 0) {
            cnt[j] = lo;
            dep[j] = up;
        }
    }
}

__global__ void k_lower_fill(int n, const int *ptr, const int *ind, const double *val, const int *lo_ptr, int *lo_row,
                             double *lo_w)
{
    const int lane = threadIdx.x & 63;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    for (int j = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; j < n; j += nw) {
        const int len = lo_ptr[j + 1] - lo_ptr[j];
        const int src = ptr[j + 1] - len, dst = lo_ptr[j];
        for (int t = lane; t < len; t += 64) {
            lo_row[dst + t] = ind[src + t];
            lo_w[dst + t] = fabs(val[src + t]);
        }
    }
}

// Plackett-Luce draw keys of the learned ordering: key = score - log(-log u),
// u = (53 Philox bits + 1/2) 2^-53; the double is mapped to an unsigned
// 64-bit word whose ascending order is the keys' descending order (split
// into high and low 32-bit halves for two stable 32-bit radix passes).
__global__ void k_gumbel_keys(int n, const float *score, uint64_t seed, uint32_t *hi, uint32_t *lo)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const U4 o = philox_keyed(seed, (uint32_t)i, 0u, kTagGumbel);
        const double u = ((double)((((uint64_t)o.y << 32) | o.x) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        const double key = (double)score[i] - log(-log(u));
        uint64_t b = (uint64_t)__double_as_longlong(key);
        b = (b >> 63) ? ~b : (b | 0x8000000000000000ull);      // ascending in key
        b = ~b;                                                  // descending in key
        hi[i] = (uint32_t)(b >> 32);
        lo[i] = (uint32_t)b;
    }
}

}  // namespace

void launch_gumbel_keys(int n, const float *score, uint64_t seed, uint32_t *hi, uint32_t *lo, hipStream_t s)
{
    hipLaunchKernelGGL(k_gumbel_keys, dim3(2048), dim3(kPB), 0, s, n, score, seed, hi, lo);
}

void launch_map_rows(int nnz, const int *ind, const int *inv, uint32_t *rowk, hipStream_t s)
{
    hipLaunchKernelGGL(k_map_rows, dim3(4096), dim3(kPB), 0, s, nnz, ind, inv, rowk);
}
void launch_gather_key(const int *src, const int *idx, uint32_t *dst, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_gather_key, dim3(4096), dim3(kPB), 0, s, src, idx, dst, n);
}
void launch_perm_finish(int nnz, const int *idx, const int *ind, const int *inv, const double *val, int *out_ind,
                        double *out_val, hipStream_t s)
{
    hipLaunchKernelGGL(k_perm_finish, dim3(4096), dim3(kPB), 0, s, nnz, idx, ind, inv, val, out_ind, out_val);
}
void launch_colsum(int n, const int *ptr, const double *val, double *cs, double *app, int *flag, hipStream_t s)
{
    hipLaunchKernelGGL(k_colsum, dim3(2048), dim3(kPB), 0, s, n, ptr, val, cs, app, flag);
}
void launch_ground_count(int n, const int *ptr, const int *flag, int extra, int *cnt, hipStream_t s)
{
    hipLaunchKernelGGL(k_ground_count, dim3(2048), dim3(kPB), 0, s, n, ptr, flag, extra, cnt);
}
void launch_ground_fill(int n, const int *ptr, const int *ind, const double *val, const double *app, const int *flag,
                        const int *fpos, double corner, const int *optr, int *oind, double *oval, hipStream_t s)
{
    hipLaunchKernelGGL(k_ground_fill, dim3(4096), dim3(kPB), 0, s, n, ptr, ind, val, app, flag, fpos, corner, optr,
                       oind, oval);
}
void launch_trim_count(int n, const int *ptr, const int *ind, int *cnt, hipStream_t s)
{
    hipLaunchKernelGGL(k_trim_count, dim3(4096), dim3(kPB), 0, s, n, ptr, ind, cnt);
}
void launch_trim_fill(int n, const int *ptr, const int *ind, const double *val, const int *optr, int *oind,
                      double *oval, hipStream_t s)
{
    hipLaunchKernelGGL(k_trim_fill, dim3(4096), dim3(kPB), 0, s, n, ptr, ind, val, optr, oind, oval);
}
void launch_lower_count(int n, const int *ptr, const int *ind, int *cnt, int *dep, hipStream_t s)
{
    hipLaunchKernelGGL(k_lower_count, dim3(4096), dim3(kPB), 0, s, n, ptr, ind, cnt, dep);
}
void launch_lower_fill(int n, const int *ptr, const int *ind, const double *val, const int *lo_ptr, int *lo_row,
                       double *lo_w, hipStream_t s)
{
    hipLaunchKernelGGL(k_lower_fill, dim3(4096), dim3(kPB), 0, s, n, ptr, ind, val, lo_ptr, lo_row, lo_w);
}

}  // namespace rc
