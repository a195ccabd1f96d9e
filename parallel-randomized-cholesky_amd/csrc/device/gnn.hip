// gnn.hip -- forward and backward kernels of the ordering network.
//
// Network (no counterpart in the reference; the reference orders offline with
// Julia heuristics, write_graph.jl:46-78): node inputs x_v = (log degree,
// log weighted degree), centred per graph; three message-passing layers
//   M_l = Anorm H_{l-1}          (Anorm[v,u] = w_vu / wdeg_v)
//   Z_l = H_{l-1} Ws_l^T + M_l Wn_l^T + b_l,   H_l = relu(Z_l)
// with hidden width 64, and a linear score s_v = H_3 wo + bo.  Weights and
// layer inputs are rounded to bfloat16, products accumulate in FP32; the
// 128 -> 64 transforms of layers 2 and 3 run on MFMA (v_mfma_f32_16x16x16
// bf16), one wave per 16 rows.  The backward pass runs in FP32 on the
// bf16-rounded operands of the forward pass (straight-through for the
// rounding); weight gradients reduce over rows with fixed per-block partials,
// so gradients are deterministic.
#include <hip/hip_bf16.h>

#include "common.cuh"
#include "kernels.h"

namespace rc {

namespace {

constexpr int kH = kGnnHidden;   // 64

__device__ __forceinline__ float bf16r(float x)
{   // round to nearest even bfloat16, returned as float
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7F800000u) == 0x7F800000u) return x;           // inf / nan unchanged
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7FFFu + lsb;
    return __uint_as_float(u & 0xFFFF0000u);
}

__device__ __forceinline__ unsigned short bf16bits(float x) { return (unsigned short)(__float_as_uint(bf16r(x)) >> 16); }

// ---------------------------------------------------------------- aggregation
// M[v, :D] = sum_u a[v,u] H[u, :D]; D = 64: one wave per row, lane = feature.
__global__ void k_agg64(int n, const int *ptr, const int *ind, const float *a, const float *H, float *M)
{
    const int lane = lane_id();
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * blockDim.x / kWave;
    for (int v = wave; v < n; v += nwaves) {
        float acc = 0.f;
        for (int e = ptr[v]; e < ptr[v + 1]; ++e) acc = fmaf(a[e], H[(size_t)ind[e] * kH + lane], acc);
        M[(size_t)v * kH + lane] = acc;
    }
}

// Transposed aggregation for the backward pass: out[u] += sum_{v in N(u)} a[v,u] G[v]
// with a[v,u] = w / wdeg_v stored at the (u -> v) entry as at[e].
__global__ void k_aggT64(int n, const int *ptr, const int *ind, const float *at, const float *G, float *out)
{
    const int lane = lane_id();
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * blockDim.x / kWave;
    for (int u = wave; u < n; u += nwaves) {
        float acc = out[(size_t)u * kH + lane];
        for (int e = ptr[u]; e < ptr[u + 1]; ++e) acc = fmaf(at[e], G[(size_t)ind[e] * kH + lane], acc);
        out[(size_t)u * kH + lane] = acc;
    }
}

__global__ void k_agg2(int n, const int *ptr, const int *ind, const float *a, const float *H, float *M)
{
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        float a0 = 0.f, a1 = 0.f;
        for (int e = ptr[v]; e < ptr[v + 1]; ++e) {
            a0 = fmaf(a[e], H[(size_t)ind[e] * 2], a0);
            a1 = fmaf(a[e], H[(size_t)ind[e] * 2 + 1], a1);
        }
        M[(size_t)v * 2] = a0;
        M[(size_t)v * 2 + 1] = a1;
    }
}

// ---------------------------------------------------------------- layer 1 (2 -> 64)
__global__ void k_dense1(int n, const float *H0, const float *M1, const float *w, float *Z, float *H)
{
    const float *Ws = w + kGnnWs1, *Wn = w + kGnnWn1, *b = w + kGnnB1;
    const int lane = lane_id();
    const float ws0 = bf16r(Ws[lane * 2]), ws1 = bf16r(Ws[lane * 2 + 1]);
    const float wn0 = bf16r(Wn[lane * 2]), wn1 = bf16r(Wn[lane * 2 + 1]);
    const float bb = b[lane];
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * blockDim.x / kWave;
    for (int v = wave; v < n; v += nwaves) {
        float h0 = bf16r(H0[(size_t)v * 2]), h1 = bf16r(H0[(size_t)v * 2 + 1]);
        float m0 = bf16r(M1[(size_t)v * 2]), m1 = bf16r(M1[(size_t)v * 2 + 1]);
        float z = bb;
        z = fmaf(ws0, h0, z);
        z = fmaf(ws1, h1, z);
        z = fmaf(wn0, m0, z);
        z = fmaf(wn1, m1, z);
        Z[(size_t)v * kH + lane] = z;
        H[(size_t)v * kH + lane] = z > 0.f ? z : 0.f;
    }
}

// ---------------------------------------------------------------- layers 2, 3 (128 -> 64) on MFMA
// Z = [H | M] W^T + b with W = [Ws | Wn] (64 x 128) in bf16.  One wave per 16
// rows: 4 output tiles of 16 x 16, K = 128 in 8 steps of 16.
// v_mfma_f32_16x16x16_bf16 operand layout (per lane l): A holds row l % 16,
// k = 4 (l / 16) .. +3; B holds column l % 16, same k; D holds rows
// 4 (l / 16) .. +3 of column l % 16.
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_dense_mfma(int n, const float *Hin, const float *Min, const float *Ws,
                                                    const float *Wn, const float *b, float *Z, float *Hout)
{
    __shared__ unsigned short sw[kH * 2 * kH];   // W[o][k], k < 128, bf16 bits
    for (int i = threadIdx.x; i < kH * 2 * kH; i += blockDim.x) {
        int o = i / (2 * kH), k = i % (2 * kH);
        float v = k < kH ? Ws[o * kH + k] : Wn[o * kH + (k - kH)];
        sw[i] = bf16bits(v);
    }
    __syncthreads();
    const int lane = lane_id();
    const int wid = threadIdx.x / kWave;
    const int r16 = lane % 16, kq = lane / 16;
    const int ntiles = (n + 15) / 16;
    for (int tile = blockIdx.x * 4 + wid; tile < ntiles; tile += gridDim.x * 4) {
        const int row = tile * 16 + r16;
        f32x4 acc[4];
        for (int c = 0; c < 4; ++c) acc[c] = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < 8; ++ks) {
            const int k0 = ks * 16 + kq * 4;
            bf16x4 av;
            for (int t = 0; t < 4; ++t) {
                int k = k0 + t;
                float x = 0.f;
                if (row < n) x = k < kH ? Hin[(size_t)row * kH + k] : Min[(size_t)row * kH + (k - kH)];
                av[t] = (short)bf16bits(x);
            }
            for (int c = 0; c < 4; ++c) {
                const int col = c * 16 + r16;
                bf16x4 bv;
                for (int t = 0; t < 4; ++t) bv[t] = (short)sw[col * 2 * kH + k0 + t];
                acc[c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[c], 0, 0, 0);
            }
        }
        for (int c = 0; c < 4; ++c) {
            const int col = c * 16 + r16;
            for (int t = 0; t < 4; ++t) {
                const int rr = tile * 16 + kq * 4 + t;
                if (rr < n) {
                    float z = acc[c][t] + b[col];
                    Z[(size_t)rr * kH + col] = z;
                    Hout[(size_t)rr * kH + col] = z > 0.f ? z : 0.f;
                }
            }
        }
    }
}

// ---------------------------------------------------------------- output
__global__ void k_score(int n, const float *H3, const float *w, float *s)
{
    const float *wo = w + kGnnWo;
    const float bo = w[kGnnBo];
    const int lane = lane_id();
    const float wl = bf16r(wo[lane]);
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * blockDim.x / kWave;
    for (int v = wave; v < n; v += nwaves) {
        float p = wl * bf16r(H3[(size_t)v * kH + lane]);
        for (int off = 32; off > 0; off >>= 1) p += __shfl_xor(p, off, kWave);
        if (lane == 0) s[v] = p + bo;
    }
}

// ---------------------------------------------------------------- backward
// dH3[v][o] = ds[v] wo[o]; partial dwo / dbo per block.
__global__ void k_score_back(int n, const float *H3, const float *w, const float *ds, float *dH, float *part)
{
    const float *wo = w + kGnnWo;
    const int lane = lane_id();
    const float wl = bf16r(wo[lane]);
    float gwo = 0.f, gbo = 0.f;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * blockDim.x / kWave;
    for (int v = wave; v < n; v += nwaves) {
        float d = ds[v];
        dH[(size_t)v * kH + lane] = d * wl;
        gwo = fmaf(d, bf16r(H3[(size_t)v * kH + lane]), gwo);
        gbo += d;
    }
    // fixed-order per-block reduction over the block's waves through LDS
    __shared__ float red[256];
    red[threadIdx.x] = gwo;
    __syncthreads();
    if (threadIdx.x < kWave) {
        float t = 0.f;
        for (int w2 = 0; w2 < (int)blockDim.x / kWave; ++w2) t += red[w2 * kWave + threadIdx.x];
        part[(size_t)blockIdx.x * (kH + 1) + threadIdx.x] = t;
    }
    __syncthreads();
    red[threadIdx.x] = lane == 0 ? gbo : 0.f;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int w2 = 0; w2 < (int)blockDim.x / kWave; ++w2) t += red[w2 * kWave];
        part[(size_t)blockIdx.x * (kH + 1) + kH] = t;
    }
}

// dZ = dH * (Z > 0) in place in dH.
__global__ void k_relu_back(size_t total, const float *Z, float *dH)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x)
        if (!(Z[i] > 0.f)) dH[i] = 0.f;
}

// Weight-gradient partials of one layer: for the rows of this block,
// gWs[o][i] = sum_v dZ[v][o] bf16(X[v][i]), gWn likewise with M, gb[o] = sum dZ.
// Thread (o, i-slice): 256 threads = 64 outputs x 4 slices of the D inputs.
template <int D>
__global__ void __launch_bounds__(256) k_wgrad(int n, int rows_per_block, const float *dZ, const float *X,
                                               const float *M, float *part)
{
    const int o = threadIdx.x % kH;
    const int slice = threadIdx.x / kH;      // 0..3
    constexpr int per = (D + 3) / 4;
    float gs[per], gn[per];
    for (int t = 0; t < per; ++t) { gs[t] = 0.f; gn[t] = 0.f; }
    float gb = 0.f;
    const int r0 = blockIdx.x * rows_per_block;
    const int r1 = min(n, r0 + rows_per_block);
    for (int v = r0; v < r1; ++v) {
        const float d = dZ[(size_t)v * kH + o];
        if (d == 0.f) continue;
        for (int t = 0; t < per; ++t) {
            const int i = slice * per + t;
            if (i < D) {
                gs[t] = fmaf(d, bf16r(X[(size_t)v * D + i]), gs[t]);
                gn[t] = fmaf(d, bf16r(M[(size_t)v * D + i]), gn[t]);
            }
        }
        gb += d;
    }
    // layout of a block's partial: [Ws (64 x D) | Wn (64 x D) | b (64)]
    float *p = part + (size_t)blockIdx.x * (2 * kH * D + kH);
    for (int t = 0; t < per; ++t) {
        const int i = slice * per + t;
        if (i < D) {
            p[o * D + i] = gs[t];
            p[kH * D + o * D + i] = gn[t];
        }
    }
    if (slice == 0) p[2 * kH * D + o] = gb;
}

// Sum block partials (fixed order) into out[0..len).
__global__ void k_sum_partials(const float *part, int nblocks, int len, float *out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < len; i += gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int b = 0; b < nblocks; ++b) s += part[(size_t)b * len + i];
        out[i] = s;
    }
}

// dX[v][i] = sum_o dZ[v][o] bf16(Ws[o][i]);  G[v][i] = sum_o dZ[v][o] bf16(Wn[o][i])
// (D = 64: one wave per row, lane = i).
__global__ void k_back_dense(int n, const float *dZ, const float *Ws, const float *Wn, float *dX, float *G)
{
    __shared__ float ws[kH * kH], wn[kH * kH];
    for (int t = threadIdx.x; t < kH * kH; t += blockDim.x) {
        ws[t] = bf16r(Ws[t]);
        wn[t] = bf16r(Wn[t]);
    }
    __syncthreads();
    const int lane = lane_id();
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * blockDim.x / kWave;
    for (int v = wave; v < n; v += nwaves) {
        const float dz = dZ[(size_t)v * kH + lane];
        float a = 0.f, g = 0.f;
        for (int o = 0; o < kH; ++o) {
            const float d = __shfl(dz, o, kWave);
            a = fmaf(d, ws[o * kH + lane], a);
            g = fmaf(d, wn[o * kH + lane], g);
        }
        dX[(size_t)v * kH + lane] = a;
        G[(size_t)v * kH + lane] = g;
    }
}

int grid_for_waves(int n)
{
    int g = (n + 3) / 4;
    if (g > 8192) g = 8192;
    return g > 0 ? g : 1;
}

}  // namespace

void gnn_forward(const GnnArgs &g, const float *w, bool keep, float *scores, hipStream_t s)
{
    const int n = g.n;
    const int gw = grid_for_waves(n);
    // layer 1
    hipLaunchKernelGGL(k_agg2, dim3(1024), dim3(256), 0, s, n, g.ptr, g.ind, g.a, g.H0, g.M1);
    hipLaunchKernelGGL(k_dense1, dim3(gw), dim3(256), 0, s, n, (const float *)g.H0, (const float *)g.M1, w, g.Z[0],
                       g.H[0]);
    // layers 2, 3
    const int gm = (((n + 15) / 16) + 3) / 4 < 4096 ? ((((n + 15) / 16) + 3) / 4 > 0 ? (((n + 15) / 16) + 3) / 4 : 1) : 4096;
    for (int l = 1; l < 3; ++l) {
        const float *Hin = g.H[keep ? l - 1 : (l - 1) & 1];
        float *Mout = g.M[keep ? l : 0];
        float *Zout = g.Z[keep ? l : 0];
        float *Hout = g.H[keep ? l : (l & 1)];
        hipLaunchKernelGGL(k_agg64, dim3(gw), dim3(256), 0, s, n, g.ptr, g.ind, g.a, Hin, Mout);
        const float *Ws = w + (l == 1 ? kGnnWs2 : kGnnWs3);
        const float *Wn = w + (l == 1 ? kGnnWn2 : kGnnWn3);
        const float *b = w + (l == 1 ? kGnnB2 : kGnnB3);
        hipLaunchKernelGGL(k_dense_mfma, dim3(gm), dim3(256), 0, s, n, Hin, (const float *)Mout, Ws, Wn, b, Zout, Hout);
    }
    const float *H3 = g.H[keep ? 2 : 0];
    hipLaunchKernelGGL(k_score, dim3(gw), dim3(256), 0, s, n, H3, w, scores);
}

void gnn_backward(const GnnArgs &g, const float *w, const float *ds, float *dw, hipStream_t s)
{
    const int n = g.n;
    const int gw = grid_for_waves(n);
    const size_t total = (size_t)n * kH;
    float *dH = g.dA;
    float *dX = g.dB;
    float *G = g.dC;
    const int nb = g.nblocks;
    const int rows = (n + nb - 1) / nb;
    // output layer
    hipLaunchKernelGGL(k_score_back, dim3(nb), dim3(256), 0, s, n, (const float *)g.H[2], w, ds, dH, g.part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(128), 0, s, (const float *)g.part, nb, kH + 1, dw + kGnnWo);
    for (int l = 2; l >= 0; --l) {
        hipLaunchKernelGGL(k_relu_back, dim3(2048), dim3(256), 0, s, total, (const float *)g.Z[l], dH);
        if (l > 0) {
            hipLaunchKernelGGL(k_wgrad<kH>, dim3(nb), dim3(256), 0, s, n, rows, (const float *)dH,
                               (const float *)g.H[l - 1], (const float *)g.M[l], g.part);
            const int len = 2 * kH * kH + kH;
            const int off = l == 1 ? kGnnWs2 : kGnnWs3;   // Ws, Wn, b are contiguous per layer
            hipLaunchKernelGGL(k_sum_partials, dim3(64), dim3(256), 0, s, (const float *)g.part, nb, len, dw + off);
            const float *Ws = w + (l == 1 ? kGnnWs2 : kGnnWs3);
            const float *Wn = w + (l == 1 ? kGnnWn2 : kGnnWn3);
            hipLaunchKernelGGL(k_back_dense, dim3(gw), dim3(256), 0, s, n, (const float *)dH, Ws, Wn, dX, G);
            hipLaunchKernelGGL(k_aggT64, dim3(gw), dim3(256), 0, s, n, g.ptr, g.ind, g.at, (const float *)G, dX);
            float *t = dH; dH = dX; dX = t;    // dH_{l-1}
        } else {
            hipLaunchKernelGGL(k_wgrad<2>, dim3(nb), dim3(256), 0, s, n, rows, (const float *)dH, (const float *)g.H0,
                               (const float *)g.M1, g.part);
            const int len = 2 * kH * 2 + kH;
            hipLaunchKernelGGL(k_sum_partials, dim3(4), dim3(256), 0, s, (const float *)g.part, nb, len, dw + kGnnWs1);
        }
    }
}

}  // namespace rc
