// FP64 detector geometry and the range-reduced phase decomposition used by the
// FP32-split kernels.  Shared by device code (kernels.hip) and host code (the
// planner in idisim.hip), so everything here is __host__ __device__.
//
// Phase models (reference): idi/simulation/cpu.py:42-66, sim.cu:43-95.
//
// Decomposition used by the FP32-split path (revolutions, 1 rev = 2 pi rad,
// krev = k / 2 pi).  With R_p = |r_p|, rh_p = r_p / R_p, u_p = rh_p - z_hat,
// s_pj = rh_p . r_j, h_pj^2 = |rh_p x r_j|^2 and D_pj = |r_p - r_j|:
//   near field:  theta/2pi = krev (R_p - z_p)                    per pixel
//                          - krev u_p . r_j + phi_j / 2pi         linear part
//                          + krev h^2 / (D + R_p - s)             curvature w_pj
//   (exact:  D - R + s = h^2 / (D + R - s), so D - z_p + z_j = (R_p - z_p) - u_p.r_j + w_pj)
//   far field:   theta/2pi = -krev u_p . r_j + phi_j / 2pi [+ krev (x^2+y^2) / (2 R_p)]
// A workgroup owns a tile of pixels with centre c.  Writing u_p = u_c + du_p:
//   tau_pj = Phi_p + A_cj + lam_p W_cj + dq_p . r'_cj
//   A_cj  = frac(-krev u_c . r_j + phi_j / 2pi)          FP64, reduced mod 1 rev
//   W_cj  = krev w_cj (nf) or krev (x^2+y^2)/(2 R_c) (secondorder)
//   lam_p = R_c / R_p,   dq_p = -krev du_p
//   r'_cj = r_j (1 + sigma_cj),  sigma_cj = (s_cj - w_cj) / D_cj  (nf), 0 (ff)
// The nf lines are the first-order expansion of w in (s, R) about the tile
// centre (dw/ds = (w - s)/D, dw/dR = -w/D); the host planner bounds the
// remainder (plan_bound in planner.hpp).  Far field (with or without
// secondorder) is exact up to rounding.  A is formed in FP64 and rounded; W,
// r', dq, lam are FP32; |dq . r'| is a few revolutions, so the FP32 phase
// keeps ~1e-7 rev.
// FP contraction is off in these FP64 helpers, so every kernel that inlines
// them forms bit-identical records (single-shot and batched results agree).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace idisim {

constexpr double kInv2Pi = 0.15915494309189533576888376337251436;
constexpr double kTwoPi = 6.28318530717958647692528676655900577;

struct Geom {
    double pix, detz, k, krev;
    double off0, off1, half0, half1;   // detoffset, n0/2 and n1/2 (true division)
    double M[9];
    int n0, n1;
    int rot, nf, scale, secondorder;
    int foldw;   // set per launch: add W_cj into A_cj (lam_p ~ 1 within the bound)
    // Extent check of the FP32-split record stage: an emitter with |r_j|^2 > ext2 (the
    // square of the extent the error planner assumed) sets *viol, a flag in host-visible
    // memory that the host reads on its next call.  viol == nullptr: no check.
    double ext2;
    unsigned int* viol;
    // Set per launch for tile variants 14, 15: refrho = rho (1 + sigma), the largest |r'_cj|, and
    // refc = T^2 / 4, T the planner's bound of the rotation angle t = (2 pi (d_k - D_k)) . r'.
    // A lane whose own bound T_k = |2 pi (d_k - D_k)| refrho exceeds T / sqrt 2 scales its pair
    // sums by 1 / (1 + T_k^2 / 2 - refc), which keeps the first-order truncation within T^2 / 4.
    float refrho, refc;
};

struct Frame {
    double R;       // |r_p|
    double rh[3];   // r_p / R
    double u[3];    // rh - z_hat, u_z formed without cancellation
    double RmZ;     // R - z_p, formed without cancellation
};

// cpu.py:76-80 at a (possibly fractional) pixel index
__host__ __device__ inline void pixel_pos(const Geom& g, double fx, double fy, double r[3])
{
#pragma clang fp contract(off)
    const double dx = (fx - g.half0) * g.pix + g.off0;
    const double dy = (fy - g.half1) * g.pix + g.off1;
    const double dz = g.detz;
    if (g.rot) {
        r[0] = g.M[0] * dx + g.M[1] * dy + g.M[2] * dz;
        r[1] = g.M[3] * dx + g.M[4] * dy + g.M[5] * dz;
        r[2] = g.M[6] * dx + g.M[7] * dy + g.M[8] * dz;
    } else {
        r[0] = dx;
        r[1] = dy;
        r[2] = dz;
    }
}

__host__ __device__ inline Frame make_frame(const double r[3])
{
#pragma clang fp contract(off)
    Frame f;
    const double rxy2 = r[0] * r[0] + r[1] * r[1];
    f.R = sqrt(rxy2 + r[2] * r[2]);
    const double iR = 1.0 / f.R;
    f.rh[0] = r[0] * iR;
    f.rh[1] = r[1] * iR;
    f.rh[2] = r[2] * iR;
    f.u[0] = f.rh[0];
    f.u[1] = f.rh[1];
    if (r[2] > 0.0) {
        f.RmZ = rxy2 / (f.R + r[2]);
        f.u[2] = -f.RmZ * iR;
    } else {
        f.RmZ = f.R - r[2];
        f.u[2] = f.rh[2] - 1.0;
    }
    return f;
}

__host__ __device__ inline Frame pixel_frame(const Geom& g, double fx, double fy)
{
    double r[3];
    pixel_pos(g, fx, fy, r);
    return make_frame(r);
}

// Half differences of the 2 x 2 pixel quad with lower corner (x, y) (fractional pixel
// indices): h_0 = (dq(x+1, y+1) - dq(x, y)) / 2 and h_1 = (dq(x, y+1) - dq(x+1, y)) / 2,
// dq = -krev (u - uc), the mirror-pair half differences of the pair tiles; with mm, also
// m_0 - m_1 = (dq(x, y) + dq(x+1, y+1) - dq(x+1, y) - dq(x, y+1)) / 2, the offset between the
// centres of the quad's two pairs.
__host__ __device__ inline void quad_half_diff(const Geom& g, const double uc[3], double x, double y,
                                               double h[2][3], double* mm = nullptr)
{
    const Frame f00 = pixel_frame(g, x, y), f11 = pixel_frame(g, x + 1.0, y + 1.0);
    const Frame f10 = pixel_frame(g, x + 1.0, y), f01 = pixel_frame(g, x, y + 1.0);
    for (int a = 0; a < 3; ++a) {
        const double q00 = -g.krev * (f00.u[a] - uc[a]), q11 = -g.krev * (f11.u[a] - uc[a]);
        const double q10 = -g.krev * (f10.u[a] - uc[a]), q01 = -g.krev * (f01.u[a] - uc[a]);
        h[0][a] = 0.5 * (q11 - q00);
        h[1][a] = 0.5 * (q01 - q10);
        if (mm != nullptr) mm[a] = 0.5 * (q00 + q11) - 0.5 * (q10 + q01);
    }
}

// Linear wave reference of tile variant 13 for the 16 x 16 pixel block centred at (px, py):
// D_k, the half differences of the quad centred there, and their central differences over
// one pixel, Jx_k = (h_k(+1, 0) - h_k(-1, 0)) / 2 and Jy_k likewise, all as FP32 values.
__host__ __device__ inline void wave_linear_ref(const Geom& g, const double uc[3], double px, double py,
                                                float D[2][3], float Jx[2][3], float Jy[2][3])
{
    double h0[2][3], hxp[2][3], hxm[2][3], hyp[2][3], hym[2][3];
    quad_half_diff(g, uc, px - 0.5, py - 0.5, h0);
    quad_half_diff(g, uc, px + 0.5, py - 0.5, hxp);
    quad_half_diff(g, uc, px - 1.5, py - 0.5, hxm);
    quad_half_diff(g, uc, px - 0.5, py + 0.5, hyp);
    quad_half_diff(g, uc, px - 0.5, py - 1.5, hym);
    for (int k = 0; k < 2; ++k) {
        for (int a = 0; a < 3; ++a) {
            D[k][a] = (float)h0[k][a];
            Jx[k][a] = (float)(0.5 * (hxp[k][a] - hxm[k][a]));
            Jy[k][a] = (float)(0.5 * (hyp[k][a] - hym[k][a]));
        }
    }
}

// Per (tile centre, emitter) record of the FP32-split kernels.  Only the large
// linear phase A needs FP64 (|k_rev u_c . r_j| reaches tens of revolutions and
// is reduced mod 1); W, sigma and the amplitude factor are small corrections
// whose relative precision only needs ~1e-6, so they are formed in FP32.
struct SplitRec {
    float x, y, z, A;   // r'_cj and the reduced linear phase
    float W, amp;       // curvature / second-order coefficient, amplitude
    float sigma, s;     // sigma_cj and s_cj (nf)
    float iD;           // 1 / D_cj (nf; used by the adjoint)
};

__host__ __device__ inline SplitRec make_split_rec(const Geom& g, const Frame& c, double ex,
                                                   double ey, double ez, double ephi, double ea)
{
#pragma clang fp contract(off)
    if (g.viol != nullptr && ex * ex + ey * ey + ez * ez > g.ext2) *g.viol = 1u;
    const double ucr = c.u[0] * ex + c.u[1] * ey + c.u[2] * ez;
    const double tau = -g.krev * ucr + ephi * kInv2Pi;
    const float fx = (float)ex, fy = (float)ey, fz = (float)ez;
    float W = 0.f, sig = 0.f, am = (float)ea, s = 0.f, iD = 0.f;
    if (g.nf) {
        s = (float)(ucr + ez);
        const float rx = (float)c.rh[0], ry = (float)c.rh[1], rz = (float)c.rh[2];
        const float hx = ry * fz - rz * fy;
        const float hy = rz * fx - rx * fz;
        const float hz = rx * fy - ry * fx;
        const float h2 = hx * hx + hy * hy + hz * hz;
        const float R = (float)c.R;
        const float a = R - s;
        const float D = sqrtf(a * a + h2);
        const float w = (a >= 0.f) ? h2 / (D + a) : D - a;
        W = (float)g.krev * w;
        sig = (s - w) / D;
        iD = 1.f / D;
        if (g.scale) am = (float)ea * (R / D);
    } else if (g.secondorder) {
        W = (float)g.krev * (fx * fx + fy * fy) / (2.f * (float)c.R);
    }
    const double tw = g.foldw ? tau + (double)W : tau;
    const double A = tw - rint(tw);
    const float f1 = 1.f + sig;
    SplitRec r;
    r.x = fx * f1;
    r.y = fy * f1;
    r.z = fz * f1;
    r.A = (float)A;
    r.W = W;
    r.amp = am;
    r.sigma = sig;
    r.s = s;
    r.iD = iD;
    return r;
}

// Work item of the persistent batched kernels: one (shot, pixel tile).
struct WorkItem {
    int32_t shot;
    int32_t tile;
    int64_t e0, e1;   // emitter range [e0, e1) in the concatenated array
};

// Pixel-tile variants.  0..3: VALU kernels, lanes form a 16 x 16 grid and each
// lane owns PX x PY pixels (16x16, 16x32, 32x32, 32x64 pixels per workgroup).
// 4: reserved (a matrix-core variant was measured 12% slower than the VALU tiles and
// is not built: the phase is a nonlinear distance, not a contraction).
// 5, 6: mirror-pair VALU kernels (one or two 2 x 2 quads per lane; 32x32 and
// 32x64 pixels per workgroup), which share one sincos per quad and emitter.
// 7: mirror pairs with a centre per pair (one quad per lane, 32x32 pixels): the
// second centre is reached by a first-order rotation, so no remainder is dropped.
// 8: four mirror pairs about a 2 x 4 block per lane (32x64 pixels), a centre per
// pair, one sincos of a centre per block.
// 9..12: the mirror pairs of 5 (9, 10) and the centred pairs of 7 (11, 12) with the
// per-pair sincos of delta_k taken from a reference: every wave owns a 16 x 16
// pixel block, the phasor of delta_k at the block centre is formed once per emitter
// and wave, and each lane rotates it to its own pair to first (9, 11) or second
// (10, 12) order.
// 13: as 9, with the rotation angle linear in the quad's offset from the block
// centre (two per-emitter slopes per pair and wave), so that a lane forms it in two
// operations instead of three.
// 14: as 11 (centred pairs, first-order rotation), with a reference for each 8 x 8
// pixel sub-block (four per wave, 16 per workgroup) instead of each wave block: the
// rotation angle is about 3/7 as large, and the lanes with the largest angles scale
// their sums so that the first-order truncation is at most T^2 / 4 instead of T^2 / 2;
// the first order then qualifies where 11 needs the second order of 12.
// 15: as 14, with the rotation between the two pair centres of a quad (the centring of 7)
// also taken from the sub-block's reference quad, formed once per emitter and sub-block
// instead of once per lane: 3 fewer FMA per quad, for a residual the planner tabulates.
constexpr int kNumTiles = 16;
constexpr int kTileNone = 4;  // reserved index, no kernel
constexpr int kTilePair1 = 5;
constexpr int kTilePair2 = 6;
constexpr int kTilePairC = 7;
constexpr int kTileBlock = 8;
constexpr int kTileRef1 = 9;
constexpr int kTileRef2 = 10;
constexpr int kTileRefC1 = 11;
constexpr int kTileRefC2 = 12;
constexpr int kTileRefL = 13;
constexpr int kTileRefS = 14;
constexpr int kTileRefSC = 15;
__host__ __device__ constexpr bool tile_pair(int v) { return v >= kTilePair1 && v <= kTileRefSC; }
__host__ __device__ constexpr bool tile_pair_centred(int v)
{
    return v == kTilePairC || v == kTileRefC1 || v == kTileRefC2 || v == kTileRefS || v == kTileRefSC;
}
// kind of wave reference: 0 none (own sincos), 1 first-order rotation, 2 second order,
// 3 first order with the angle linear in the quad offset, 4 first order from a reference
// per 8 x 8 pixel sub-block (the far lanes' sums scaled), 5 as 4 with the pair-centre
// rotation from the sub-block reference
__host__ __device__ constexpr int tile_ref_kind(int v)
{
    return (v == kTileRef1 || v == kTileRefC1) ? 1
           : ((v == kTileRef2 || v == kTileRefC2) ? 2
                                                  : (v == kTileRefL ? 3 : (v == kTileRefS ? 4 : (v == kTileRefSC ? 5 : 0))));
}
// order of the rotation from the reference to the lane's delta_k (0: own sincos)
__host__ __device__ constexpr int tile_ref_order(int v)
{
    return tile_ref_kind(v) >= 3 ? 1 : tile_ref_kind(v);
}
__host__ __device__ constexpr int pair_quads(int v) { return v == kTilePair2 ? 2 : 1; }
__host__ __device__ constexpr int tile_px(int v) { return v >= 2 && v != kTileNone ? 2 : 1; }
__host__ __device__ constexpr int tile_py(int v)
{
    return v == 0 || v == kTileNone ? 1 : (v == 3 || v == kTilePair2 || v == kTileBlock ? 4 : 2);
}
__host__ __device__ constexpr int tile_rows(int v) { return 16 * tile_px(v); }
__host__ __device__ constexpr int tile_cols(int v) { return 16 * tile_py(v); }

constexpr int kBlock = 256;   // threads per workgroup (4 wave64)
constexpr int kChunk = 256;   // emitters staged in LDS per round
constexpr int kSplitGrain = 64;  // finest emitter split over workgroups

// Emitters per workgroup when n emitters are split over `splits` workgroups: whole
// LDS chunks when every split gets at least one (a trailing partial chunk measured
// slower), else multiples of 64 so that small n still spreads over many workgroups.
__host__ __device__ inline int64_t split_size(int64_t n, int splits)
{
    const int64_t each = (n + splits - 1) / splits;
    const int64_t grain = each >= kChunk ? kChunk : kSplitGrain;
    const int64_t per = (each + grain - 1) / grain * grain;
    return per > 0 ? per : grain;
}

}  // namespace idisim
