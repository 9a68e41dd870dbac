// sin and cos of an FP64 phase (radians) for the plain FP64 kernels (no reference counterpart:
// the reference calls libm's and CUDA's sincos, cpu.py:50-55 and sim.cu:56-64).
//
// x = n pi/2 + r by Cody and Waite's reduction: pi/2 is split into C1 and C2 of 27 significant
// bits each, so that n C1 and n C2 are exact for |n| <= 2^26, and the rounded remainder C3 (the
// three parts from a 300-bit pi).  r = ((x - n C1) - n C2) - n C3 in FMAs is within an ulp of r
// plus 2^26 ulp(C3) ~ 1e-24.  On |r| <= pi/4 (a little more where x n 2/pi rounds the other way)
// the Taylor series of sin r to r^17 and of cos r to r^16 leave out less than (pi/4)^19 / 19! and
// (pi/4)^18 / 18!, both below 1e-19; they are evaluated in r^2 by Horner's rule with FMAs, the
// coefficients 1/k! rounded once.  The quadrant n mod 4 swaps and negates.  Beyond
// |x| = 2^26 C1 (1.05e8 rad) the library's sincos.
//
// The reference formula's near-field phase, k (|r_p - r_j| - z_p + z_j) + phi_j, stays below
// 1e8 rad on detectors up to about 0.1 rad from the axis at k = 4e4 / um; the far-field phase
// k |u_p| |r_j| is a few hundred rad.
#pragma once

#include <cmath>

namespace idisim {

// a * b + c with c a constant: v_fma_f64 with the constant in an SGPR pair.  The compiler's
// choice for a Horner step is v_fmac_f64, whose destination is its addend, so every step copied
// the constant into a fresh register first (one v_mov_b64 per step).
__host__ __device__ inline double fma_const(double a, double b, double c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
    return d;
#else
    return fma(a, b, c);
#endif
}

__host__ __device__ inline void sincos_fp64(double x, double* s, double* c)
{
    constexpr double kC1 = 0x1.921fb54p+0;
    constexpr double kC2 = 0x1.10b461p-30;
    constexpr double kC3 = 0x1.a62633145c06ep-58;
    constexpr double k2OverPi = 0x1.45f306dc9c883p-1;
    constexpr double kLim = 0x1.921fb54p+26;  // 2^26 C1
    if (!(fabs(x) < kLim)) {
        sincos(x, s, c);
        return;
    }
    const double n = rint(x * k2OverPi);
    double r = fma(-n, kC1, x);
    r = fma(-n, kC2, r);
    r = fma(-n, kC3, r);
    const double r2 = r * r;
    // sin r = r + r^3 (-1/3! + r^2 (1/5! + r^2 (... + r^2 / 17!)))
    double ps = 1.0 / 355687428096000.0;
    ps = fma_const(ps, r2, -1.0 / 1307674368000.0);
    ps = fma_const(ps, r2, 1.0 / 6227020800.0);
    ps = fma_const(ps, r2, -1.0 / 39916800.0);
    ps = fma_const(ps, r2, 1.0 / 362880.0);
    ps = fma_const(ps, r2, -1.0 / 5040.0);
    ps = fma_const(ps, r2, 1.0 / 120.0);
    ps = fma_const(ps, r2, -1.0 / 6.0);
    const double sr = fma(r * r2, ps, r);
    // cos r = 1 + r^2 (-1/2 + r^2 (1/4! + r^2 (... + r^2 / 16!)))
    double pc = 1.0 / 20922789888000.0;
    pc = fma_const(pc, r2, -1.0 / 87178291200.0);
    pc = fma_const(pc, r2, 1.0 / 479001600.0);
    pc = fma_const(pc, r2, -1.0 / 3628800.0);
    pc = fma_const(pc, r2, 1.0 / 40320.0);
    pc = fma_const(pc, r2, -1.0 / 720.0);
    pc = fma_const(pc, r2, 1.0 / 24.0);
    pc = fma_const(pc, r2, -0.5);
    const double cr = fma(r2, pc, 1.0);
    // x = n pi/2 + r: quadrant q = n mod 4 (two's complement, also for n < 0)
    const int q = (int)n & 3;
    const double sa = (q & 1) ? cr : sr, ca = (q & 1) ? sr : cr;
    *s = (q & 2) ? -sa : sa;
    *c = ((q + 1) & 2) ? -ca : ca;
}

}  // namespace idisim
