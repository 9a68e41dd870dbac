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
// Many balls (simobj.multisphere, simobj.py:230-262): emitter i is drawn as in
// one ball and shifted by the centre of its ball (host-side centres and counts).
// Gaussian cloud (simobj.gauss / gnorm with rho = 2, simobj.py:154-184): normal
// coordinates by Box-Muller, scaled per axis and optionally rotated.
#include "kernels.hpp"
#include "philox.hpp"

namespace idisim {

namespace {
// Uniform point in the unit-radius ball (times `radius`) and a phase in [0, 2 pi) for
// emitter i of `stream`: blocks (i, stream, 0) and (i, stream, 1).
__device__ inline void ball_draw(uint64_t i, uint64_t stream, uint32_t k0, uint32_t k1, double radius, double* p,
                                 double* phase)
{
    const U4 c0 = {{(uint32_t)i, (uint32_t)(i >> 32), (uint32_t)stream, 0u}};
    const U4 c1 = {{(uint32_t)i, (uint32_t)(i >> 32), (uint32_t)stream, 1u}};
    const U4 r0 = philox4x32_10(c0, k0, k1);
    const U4 r1 = philox4x32_10(c1, k0, k1);
    const double t = 2.0 * u53(r0.v[0], r0.v[1]) - 1.0;
    const double az = kTwoPi * u53(r0.v[2], r0.v[3]);
    const double rr = radius * cbrt(u53(r1.v[0], r1.v[1]));
    const double st = sqrt(1.0 - t * t);
    double sa, ca;
    sincos(az, &sa, &ca);
    p[0] = rr * t;
    p[1] = rr * st * sa;
    p[2] = rr * st * ca;
    *phase = kTwoPi * u53(r1.v[2], r1.v[3]);
}

// Box-Muller pair from one block: r = sqrt(-2 ln u1) (u1 in (0, 1]), angle 2 pi u2.
__device__ inline void normal_pair(const U4& r, double* z0, double* z1)
{
    const double u1 = 1.0 - u53(r.v[0], r.v[1]);
    const double rad = sqrt(-2.0 * log(u1));
    double sa, ca;
    sincos(kTwoPi * u53(r.v[2], r.v[3]), &sa, &ca);
    *z0 = rad * ca;
    *z1 = rad * sa;
}
}  // namespace

__global__ void k_gen_ball(uint64_t seed, uint64_t stream, int64_t n, double radius, double cx, double cy, double cz,
                           int positions, int phases, double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double p[3], ph;
    ball_draw((uint64_t)i, stream, (uint32_t)seed, (uint32_t)(seed >> 32), radius, p, &ph);
    double* o = out + 4 * i;
    if (positions) {
        o[0] = cx + p[0];
        o[1] = cy + p[1];
        o[2] = cz + p[2];
    }
    if (phases) o[3] = ph;
}

// Emitter i belongs to ball b with offsets[b] <= i < offsets[b + 1] (binary search,
// clamped to a valid ball for any offsets).
__global__ void k_gen_balls(uint64_t seed, uint64_t stream, int64_t n, double radius, int64_t nballs,
                            const double* __restrict__ centres, const int64_t* __restrict__ offsets, int positions,
                            int phases, double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double p[3], ph;
    ball_draw((uint64_t)i, stream, (uint32_t)seed, (uint32_t)(seed >> 32), radius, p, &ph);
    double* o = out + 4 * i;
    if (positions) {
        int64_t lo = 0, hi = nballs - 1;  // last b with offsets[b] <= i
        while (lo < hi) {
            const int64_t mid = lo + (hi - lo + 1) / 2;
            if (offsets[mid] <= i)
                lo = mid;
            else
                hi = mid - 1;
        }
        o[0] = centres[3 * lo] + p[0];
        o[1] = centres[3 * lo + 1] + p[1];
        o[2] = centres[3 * lo + 2] + p[2];
    }
    if (phases) o[3] = ph;
}

// Gaussian cloud: (sx z0, sy z1, sz z2) with z from blocks (i, stream, 0) and (i, stream, 1),
// optionally rotated, p -> M p; phase from block (i, stream, 2).
__global__ void k_gen_gauss(uint64_t seed, uint64_t stream, int64_t n, double sx, double sy, double sz, Rot3 M,
                            int rotate, int positions, int phases, double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const uint64_t u = (uint64_t)i;
    double* o = out + 4 * i;
    if (positions) {
        const U4 c0 = {{(uint32_t)u, (uint32_t)(u >> 32), (uint32_t)stream, 0u}};
        const U4 c1 = {{(uint32_t)u, (uint32_t)(u >> 32), (uint32_t)stream, 1u}};
        double z0, z1, z2, unused;
        normal_pair(philox4x32_10(c0, k0, k1), &z0, &z1);
        normal_pair(philox4x32_10(c1, k0, k1), &z2, &unused);
        const double x = sx * z0, y = sy * z1, z = sz * z2;
        if (rotate) {
            o[0] = M.m[0] * x + M.m[1] * y + M.m[2] * z;
            o[1] = M.m[3] * x + M.m[4] * y + M.m[5] * z;
            o[2] = M.m[6] * x + M.m[7] * y + M.m[8] * z;
        } else {
            o[0] = x;
            o[1] = y;
            o[2] = z;
        }
    }
    if (phases) {
        const U4 c2 = {{(uint32_t)u, (uint32_t)(u >> 32), (uint32_t)stream, 2u}};
        const U4 r2 = philox4x32_10(c2, k0, k1);
        o[3] = kTwoPi * u53(r2.v[0], r2.v[1]);
    }
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

hipError_t launch_gen_balls(uint64_t seed, uint64_t stream, int64_t n, double radius, int64_t nballs,
                            const double* centres, const int64_t* offsets, int positions, int phases, double* out,
                            hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gen_balls, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, seed, stream, n, radius,
                       nballs, centres, offsets, positions, phases, out);
    return hipGetLastError();
}

hipError_t launch_gen_gauss(uint64_t seed, uint64_t stream, int64_t n, const double* sigma, const double* rot,
                            int positions, int phases, double* out, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    Rot3 M{};
    if (rot)
        for (int k = 0; k < 9; ++k) M.m[k] = rot[k];
    hipLaunchKernelGGL(k_gen_gauss, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, seed, stream, n, sigma[0],
                       sigma[1], sigma[2], M, rot ? 1 : 0, positions, phases, out);
    return hipGetLastError();
}

}  // namespace idisim
