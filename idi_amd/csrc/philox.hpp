// Philox4x32-10, the counter-based generator of Salmon, Moraes, Dror and Shaw,
// "Parallel random numbers: as easy as 1, 2, 3" (SC'11): 10 rounds of
//   (c0, c1, c2, c3) -> (hi(M1 c2) ^ c1 ^ k0, lo(M1 c2), hi(M0 c0) ^ c3 ^ k1, lo(M0 c0))
// with M0 = 0xD2511F53, M1 = 0xCD9E8D57 and the key bumped by (0x9E3779B9,
// 0xBB67AE85) between rounds.  Any (counter, key) block can be drawn
// independently, so every emitter / photon hit has its own numbers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace idisim {

struct U4 {
    uint32_t v[4];
};

__host__ __device__ inline U4 philox4x32_10(U4 ctr, uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * ctr.v[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * ctr.v[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        U4 n;
        n.v[0] = hi1 ^ ctr.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ ctr.v[3] ^ k1;
        n.v[3] = lo0;
        ctr = n;
    }
    return ctr;
}

// 53-bit double in [0, 1) from two 32-bit words (27 + 26 bits).
__host__ __device__ inline double u53(uint32_t a, uint32_t b)
{
    return (double)(((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6)) * (1.0 / 9007199254740992.0);
}

}  // namespace idisim
