// On-device emitter generation (no reference counterpart; it moves the host
// draws of simobj.sphere (simobj.py:119-130 -> util/random.py:39-51) and of
// the random phases (simobj.py:56-57) onto the GPU).
//
// Random numbers: Philox4x32-10, the counter-based generator of Salmon,
// Moraes, Dror and Shaw, "Parallel random numbers: as easy as 1, 2, 3"
// (SC'11).  Emitter i of stream s draws two blocks with counters
// (i_lo, i_hi, s, 0) and (i_lo, i_hi, s, 1) under key (seed_lo, seed_hi), so
// any emitter of any shot can be regenerated independently.  Each double in
// [0, 1) takes 53 bits from two 32-bit words (27 + 26 bits).
// Ball sampling (same construction as util/random.py:39-51): t = 2u0 - 1,
// azimuth 2 pi u1, radius R u2^(1/3); phase 2 pi u3.
#include "kernels.hpp"

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

__host__ __device__ inline double u53(uint32_t a, uint32_t b)
{
    return (double)(((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6)) * (1.0 / 9007199254740992.0);
}

__global__ void k_gen_ball(uint64_t seed, uint64_t stream, int64_t n, double radius, double cx, double cy, double cz,
                           int positions, int phases, double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    U4 c0 = {{(uint32_t)i, (uint32_t)((uint64_t)i >> 32), (uint32_t)stream, 0u}};
    U4 c1 = {{(uint32_t)i, (uint32_t)((uint64_t)i >> 32), (uint32_t)stream, 1u}};
    const U4 r0 = philox4x32_10(c0, k0, k1);
    const U4 r1 = philox4x32_10(c1, k0, k1);
    double* o = out + 4 * i;
    if (positions) {
        const double t = 2.0 * u53(r0.v[0], r0.v[1]) - 1.0;
        const double az = kTwoPi * u53(r0.v[2], r0.v[3]);
        const double rr = radius * cbrt(u53(r1.v[0], r1.v[1]));
        const double st = sqrt(1.0 - t * t);
        double sa, ca;
        sincos(az, &sa, &ca);
        o[0] = cx + rr * t;
        o[1] = cy + rr * st * sa;
        o[2] = cz + rr * st * ca;
    }
    if (phases) o[3] = kTwoPi * u53(r1.v[2], r1.v[3]);
}

hipError_t launch_gen_ball(uint64_t seed, uint64_t stream, int64_t n, double radius, const double* centre,
                           int positions, int phases, double* out, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const double cx = centre ? centre[0] : 0.0, cy = centre ? centre[1] : 0.0, cz = centre ? centre[2] : 0.0;
    hipLaunchKernelGGL(k_gen_ball, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, seed, stream, n, radius, cx,
                       cy, cz, positions, phases, out);
    return hipGetLastError();
}

}  // namespace idisim
