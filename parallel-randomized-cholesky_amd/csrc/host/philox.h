// philox.h -- Philox4x32-10 counter-based generator (Salmon, Moraes, Dror,
// Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), usable from
// host and device code.  Every random draw of the library is a pure function
// of (seed, counter), so the elimination is independent of the schedule.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RC_HD __host__ __device__ __forceinline__
#else
#define RC_HD inline
#endif

namespace rc {

// Counter word 2 separates the consumers of random numbers.
enum : uint32_t { kTagSample = 0, kTagNormal = 1, kTagPerm = 2, kTagWeight = 3, kTagPoint = 4, kTagGumbel = 5 };

struct U4 { uint32_t x, y, z, w; };

RC_HD U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += W0; k1 += W1; }
        uint64_t p0 = (uint64_t)M0 * c0;
        uint64_t p1 = (uint64_t)M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    return U4{c0, c1, c2, c3};
}

RC_HD U4 philox_keyed(uint64_t seed, uint32_t a, uint32_t b, uint32_t tag)
{
    return philox4x32_10(a, b, tag, 0u, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// (0, 1]: 53 random bits plus one, times 2^-53.
RC_HD double u01_open0(uint32_t lo, uint32_t hi)
{
    uint64_t bits = ((((uint64_t)hi) << 32) | lo) >> 11;
    return (double)(bits + 1) * 0x1.0p-53;
}
// [0, 1)
RC_HD double u01_open1(uint32_t lo, uint32_t hi)
{
    uint64_t bits = ((((uint64_t)hi) << 32) | lo) >> 11;
    return (double)bits * 0x1.0p-53;
}

// Sampling draw of the elimination: entry i of column k.
RC_HD double sample_uniform(uint64_t seed, uint32_t i, uint32_t k)
{
    U4 o = philox_keyed(seed, i, k, kTagSample);
    return u01_open0(o.x, o.y);
}

}  // namespace rc
