// On-device emitter generation (no reference counterpart; it moves the host
// draws of simobj.sphere (simobj.py:119-130 -> util/random.py:39-51) and of
// the random phases (simobj.py:56-57) onto the GPU).
//
// Random numbers: Philox4x32-10 (philox.hpp).  Emitter i of stream s draws two blocks with counters
// (i_lo, i_hi, s, 0) and (i_lo, i_hi, s, 1) under key (seed_lo, seed_hi), so
// any emitter of any shot can be regenerated independently.  Each double in
// [0, 1) takes 53 bits from two 32-bit words (27 + 26 bits).
// Ball sampling (same construction as util/random.py:39-51): t = 2u0 - 1,
// azimuth 2 pi u1, radius R u2^(1/3); phase 2 pi u3.
#include "kernels.hpp"
#include "philox.hpp"

namespace idisim {

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
