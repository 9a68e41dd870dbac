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

// (cos, sin) of the per-pixel near-field phase 2 pi krev (R_p - z_p), range-reduced in FP64.
__device__ inline void pixel_phasor(const Geom& g, const Frame& f, double& cs, double& sn)
{
    double ph = g.krev * f.RmZ;
    ph -= rint(ph);
    sincos(kTwoPi * ph, &sn, &cs);
}

// Final per-pixel step of the FP32-split path: the per-pixel phase
// krev (R_p - z_p) of the near-field model (dropped by |E|^2, restored for the
// field) and the 1/|r_p| factor of "scale".  split_field is the complex64 field value, given
// the pixel's frame f and (nf) its pixel_phasor.
__device__ inline float2 split_field(const Geom& g, const Frame& f, double cs, double sn, double re, double im)
{
    if (g.scale) {
        const double iR = 1.0 / f.R;
        re *= iR;
        im *= iR;
    }
    if (g.nf) {
        const double r2 = re * cs - im * sn;
        const double i2 = re * sn + im * cs;
        re = r2;
        im = i2;
    }
    return make_float2((float)re, (float)im);
}

__device__ inline void store_split(const Geom& g, int x, int y, double re, double im, void* out,
                                   int kind, size_t idx)
{
    if (kind == OUT_INTENSITY && !g.scale) {
        static_cast<float*>(out)[idx] = (float)(re * re + im * im);
        return;
    }
    if (kind == OUT_INTENSITY) {
        const Frame f = pixel_frame(g, (double)x, (double)y);
        const double iR = 1.0 / f.R;
        re *= iR;
        im *= iR;
        static_cast<float*>(out)[idx] = (float)(re * re + im * im);
        return;
    }
    const Frame f = pixel_frame(g, (double)x, (double)y);
    double cs = 1.0, sn = 0.0;
    if (g.nf) pixel_phasor(g, f, cs, sn);
    static_cast<float2*>(out)[idx] = split_field(g, f, cs, sn, re, im);
}

// The FP64 sum over the splits of pixel idx's float4 partials (hi - lo of re and im), in the
// order of k_finalize_split: four rows of splits s = r mod 4, each summed in two halves (s = r
// mod 8 and s = r + 4 mod 8), then (r0 + r1) + (r2 + r3).  The same roundings, so that a kernel
// that sums the partials itself writes the same bits as k_finalize_split.
__device__ inline void split_sum(const float4* __restrict__ partial, int splits, size_t npix, size_t idx, double& re,
                                 double& im)
{
    double rr[4], ri[4];
    if (splits <= 32) {
        // every partial in flight at once (loads past the last split read the last one, unused),
        // then the sums of the loop below in its order
        float4 v[32];
#pragma unroll
        for (int s = 0; s < 32; ++s) v[s] = partial[(size_t)min(s, splits - 1) * npix + idx];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            double re0 = 0.0, im0 = 0.0, re1 = 0.0, im1 = 0.0;
            bool go = true;
#pragma unroll
            for (int s = r; s < 32; s += 8) {
                const bool two = go && s + 4 < splits;
                if (two) {
                    re0 += (double)v[s].x - (double)v[s].z;
                    im0 += (double)v[s].y - (double)v[s].w;
                    re1 += (double)v[s + 4].x - (double)v[s + 4].z;
                    im1 += (double)v[s + 4].y - (double)v[s + 4].w;
                } else if (go && s < splits) {
                    re0 += (double)v[s].x - (double)v[s].z;
                    im0 += (double)v[s].y - (double)v[s].w;
                }
                go = two;
            }
            rr[r] = re0 + re1;
            ri[r] = im0 + im1;
        }
        re = (rr[0] + rr[1]) + (rr[2] + rr[3]);
        im = (ri[0] + ri[1]) + (ri[2] + ri[3]);
        return;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        double re0 = 0.0, im0 = 0.0, re1 = 0.0, im1 = 0.0;
        int s = r;
        for (; s + 4 < splits; s += 8) {
            const float4 a = partial[(size_t)s * npix + idx];
            const float4 b = partial[(size_t)(s + 4) * npix + idx];
            re0 += (double)a.x - (double)a.z;
            im0 += (double)a.y - (double)a.w;
            re1 += (double)b.x - (double)b.z;
            im1 += (double)b.y - (double)b.w;
        }
        if (s < splits) {
            const float4 a = partial[(size_t)s * npix + idx];
            re0 += (double)a.x - (double)a.z;
            im0 += (double)a.y - (double)a.w;
        }
        rr[r] = re0 + re1;
        ri[r] = im0 + im1;
    }
    re = (rr[0] + rr[1]) + (rr[2] + rr[3]);
    im = (ri[0] + ri[1]) + (ri[2] + ri[3]);
}

__device__ inline void store_fp64(int kind, double re, double im, void* out, size_t idx)
{
    if (kind == OUT_INTENSITY)
        static_cast<double*>(out)[idx] = re * re + im * im;
    else
        static_cast<double2*>(out)[idx] = make_double2(re, im);
}

}  // namespace idisim
