// common.cuh -- device helpers shared by the elimination, assembly and PCG
// kernels: Philox4x32-10, agent-scope atomics, wave64 primitives, sentinels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../host/philox.h"

namespace rc {

constexpr int kWave = 64;

// ---------------------------------------------------------------- atomics
// All cross-workgroup communication goes through agent-scope atomics.  On
// gfx950 relaxed agent loads/stores are single `sc1` accesses; release and
// acquire fences add the L2 write-back / invalidate.
template <typename T>
__device__ __forceinline__ T ld_relaxed(const T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_relaxed(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T fetch_add(T *p, T v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T fetch_max(T *p, T v)
{
    return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T exchange(T *p, T v)
{
    return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void fence_release() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }
__device__ __forceinline__ void fence_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
__device__ __forceinline__ void fence_acq_rel() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent"); }

// ---------------------------------------------------------------- wave ops
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int lanes_below(uint64_t mask)
{
    return __popcll(mask & ((1ull << lane_id()) - 1ull));
}

template <typename T>
__device__ __forceinline__ T bcast(T v, int src) { return __shfl(v, src, kWave); }

__device__ __forceinline__ double wave_sum(double v)
{   // fixed butterfly order: deterministic for a fixed lane assignment
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}
__device__ __forceinline__ int wave_sum_int(int v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}
__device__ __forceinline__ int64_t wave_max(int64_t v)
{
    for (int off = 32; off > 0; off >>= 1) {
        int64_t o = __shfl_xor(v, off, kWave);
        v = o > v ? o : v;
    }
    return v;
}

// ---------------------------------------------------------------- sentinels
// Triangular-solve outputs are initialised to a signalling-NaN pattern that
// arithmetic never produces; a reader that sees it waits.  A computed value
// that happens to carry the pattern is replaced by the default quiet NaN.
constexpr uint64_t kSent64 = 0x7FF4A5A5DEADBEEFull;
constexpr uint32_t kSent32 = 0x7FA5A5A5u;

__device__ __forceinline__ uint64_t bits_of(double v) { return __double_as_longlong(v); }
__device__ __forceinline__ uint32_t bits_of(float v) { return __float_as_uint(v); }
__device__ __forceinline__ bool is_sent(double v) { return bits_of(v) == kSent64; }
__device__ __forceinline__ bool is_sent(float v) { return bits_of(v) == kSent32; }
__device__ __forceinline__ double sent_value(double) { return __longlong_as_double((long long)kSent64); }
__device__ __forceinline__ float sent_value(float) { return __uint_as_float(kSent32); }
__device__ __forceinline__ double unsent(double v) { return is_sent(v) ? __longlong_as_double(0x7FF8000000000000ll) : v; }
__device__ __forceinline__ float unsent(float v) { return is_sent(v) ? __uint_as_float(0x7FC00000u) : v; }

__device__ __forceinline__ double ld_relaxed_val(const double *p)
{
    return __longlong_as_double((long long)ld_relaxed((const unsigned long long *)p));
}
__device__ __forceinline__ float ld_relaxed_val(const float *p)
{
    return __uint_as_float(ld_relaxed((const unsigned int *)p));
}
__device__ __forceinline__ void st_relaxed_val(double *p, double v)
{
    st_relaxed((unsigned long long *)p, (unsigned long long)__double_as_longlong(v));
}
__device__ __forceinline__ void st_relaxed_val(float *p, float v)
{
    st_relaxed((unsigned int *)p, __float_as_uint(v));
}

__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Bounded waits: a wave that waits longer than this many wall-clock ticks
// (100 MHz constant clock) gives up, raises kStatusTimeout and exits.
constexpr long long kSpinTicks = 2000000000ll;   // 20 s

}  // namespace rc
