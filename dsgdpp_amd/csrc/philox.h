// Counter-based random numbers shared by host code and gfx950 kernels.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11).  A stream is a 64-bit key; draw k of a stream is the
// 64-bit half (k & 1) of Philox(counter = {k >> 1 as 64 bits, 0, 0}, key).
// nextInt(b) = floor(u64 * b / 2^64).  Because draw k is a pure function of
// (key, k), a kernel can produce any draw of a task's stream without replaying
// the ones before it -- the host only hands over the first draw index of each
// block.
//
// This generator stands in for rg::Random32 of the reference (an un-vendored
// dependency); the reference's draw *structure* is kept: one master stream per
// seed (tools/mfdsgdpp.cc:542), one stream per task seeded from the master
// (mf/sgd/dsgdpp_impl.h:79,344).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define DSGDPP_HD __host__ __device__ __forceinline__
#else
#define DSGDPP_HD inline
#endif

namespace dsgdpp {

DSGDPP_HD void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// Draw k of the stream with key (k0, k1).
DSGDPP_HD uint64_t stream_draw(uint32_t k0, uint32_t k1, uint64_t k) {
    uint32_t c[4] = {(uint32_t)(k >> 1), (uint32_t)(k >> 33), 0u, 0u};
    philox4x32_10(c, k0, k1);
    return (k & 1) ? ((uint64_t)c[2] | ((uint64_t)c[3] << 32))
                   : ((uint64_t)c[0] | ((uint64_t)c[1] << 32));
}

DSGDPP_HD uint64_t bounded(uint64_t u, uint64_t bound) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(u, bound);
#else
    return (uint64_t)(((unsigned __int128)u * bound) >> 64);
#endif
}

// Key of the master stream for a run seed, and of a task stream for the seed
// the master handed to that task.
DSGDPP_HD void master_key(uint64_t seed, uint32_t* k0, uint32_t* k1) {
    *k0 = (uint32_t)seed;
    *k1 = (uint32_t)(seed >> 32) ^ 0xD5A61266u;
}
DSGDPP_HD void task_key(uint64_t s, uint32_t* k0, uint32_t* k1) {
    *k0 = (uint32_t)s;
    *k1 = (uint32_t)(s >> 32);
}

}  // namespace dsgdpp
