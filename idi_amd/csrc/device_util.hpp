// Device helpers shared by the forward kernels (kernels.hip).
#pragma once
#include "kernels.hpp"

namespace idisim {

__device__ __forceinline__ void kahan_add(float& sum, float& comp, float v)
{
    const float y = v - comp;
    const float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

__device__ __forceinline__ void kahan_add_d(double& sum, double& comp, double v)
{
    const double y = v - comp;
    const double t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

// Pixel-index centre of the part of a tile that lies on the detector.
__device__ __forceinline__ void tile_centre(const Geom& g, int tx0, int ty0, int TX, int TY,
                                            double& cx, double& cy)
{
    const int xe = min(tx0 + TX, g.n0) - 1;
    const int ye = min(ty0 + TY, g.n1) - 1;
    cx = 0.5 * (double)(tx0 + xe);
    cy = 0.5 * (double)(ty0 + ye);
}

// Final per-pixel step of the FP32-split path: the per-pixel phase
// krev (R_p - z_p) of the near-field model (dropped by |E|^2, restored for the
// field) and the 1/|r_p| factor of "scale".
__device__ inline void store_split(const Geom& g, int x, int y, double re, double im, void* out,
                                   int kind, size_t idx)
{
    if (kind == OUT_INTENSITY && !g.scale) {
        static_cast<float*>(out)[idx] = (float)(re * re + im * im);
        return;
    }
    const Frame f = pixel_frame(g, (double)x, (double)y);
    if (g.scale) {
        const double iR = 1.0 / f.R;
        re *= iR;
        im *= iR;
    }
    if (kind == OUT_INTENSITY) {
        static_cast<float*>(out)[idx] = (float)(re * re + im * im);
        return;
    }
    if (g.nf) {
        double ph = g.krev * f.RmZ;
        ph -= rint(ph);
        double sn, cs;
        sincos(kTwoPi * ph, &sn, &cs);
        const double r2 = re * cs - im * sn;
        const double i2 = re * sn + im * cs;
        re = r2;
        im = i2;
    }
    static_cast<float2*>(out)[idx] = make_float2((float)re, (float)im);
}

__device__ inline void store_fp64(int kind, double re, double im, void* out, size_t idx)
{
    if (kind == OUT_INTENSITY)
        static_cast<double*>(out)[idx] = re * re + im * im;
    else
        static_cast<double2*>(out)[idx] = make_double2(re, im);
}

}  // namespace idisim
