// The host-side error planner of the FP32-split kernels: for every tile variant, an a-priori
// bound of the per-term phase error (revolutions) from the detector geometry and the emitter
// extent rho = max |r_j|.  The sum of: the FP32 roundings of the phase, the Taylor remainder of
// the near-field expansion about the tile centre, the fold of the curvature into the reduced
// phase, the measured v_sin_f32 / v_cos_f32 error and, for the mirror-pair tiles, the remainders
// of their shared sincos (the pair remainder, the rotation between pair centres, the rotation
// from a per-wave reference).  Tiles whose bound exceeds the tolerance are not used.
//
// Host code only (idisim.hip, and a sanitizer build of the planner tests compiled with a host
// compiler); geometry.hpp supplies the FP64 frames the kernels use.
#pragma once
#include "geometry.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

namespace idisim {

// Largest error of v_sin_f32 / v_cos_f32 as an equivalent phase: measured on
// MI355X against FP64 sin/cos over 4M arguments in [-64, 64] rev, max |dz| =
// 1.29e-7 (2.05e-8 rev); with margin.
constexpr double kSinCosErrRev = 3.0e-8;

// Planner state of a handle: the geometry, the options that select tiles and the per-geometry
// tables that planner_tables fills.
struct Planner {
    Geom g{};
    double tol_rev = 5e-6;  // per-term phase tolerance
    int summode = 0;        // 1: Kahan on every term (the mirror-pair tiles have no such mode)
    double du_max[kNumTiles] = {};
    double dR_max[kNumTiles] = {};
    double Rmin = 0.0, Rmax = 0.0;
    double pair_coef = 0.0;   // mirror-pair remainder per unit emitter extent (rev per length)
    double block_coef = 0.0;  // bound of |m_k - m_1| of the 2 x 4 block pairs per unit extent
    double ref_coef = 0.0;    // bound of |d_k - D_k| (lane pair vs wave reference) per unit extent
    double refs_coef = 0.0;   // the same against the 8 x 8 pixel sub-block reference (variants 14, 15)
    double cent_coef = 0.0;   // largest |(m_0 - m_1) - (m_0 - m_1) of the sub-block reference| (variant 15)
    double lin_coef = 0.0;    // largest residual of the linear wave reference (variant 13) per unit extent
};

inline void planner_tables(Planner* h)
{
    const Geom& g = h->g;
    h->Rmin = std::numeric_limits<double>::infinity();
    h->Rmax = 0.0;
    for (int v = 0; v < kNumTiles; ++v) {
        const int TX = tile_rows(v), TY = tile_cols(v);
        double du = 0.0, dR = 0.0;
        for (int tx0 = 0; tx0 < g.n0; tx0 += TX) {
            for (int ty0 = 0; ty0 < g.n1; ty0 += TY) {
                const int xe = std::min(tx0 + TX, g.n0) - 1, ye = std::min(ty0 + TY, g.n1) - 1;
                const Frame c = pixel_frame(g, 0.5 * (tx0 + xe), 0.5 * (ty0 + ye));
                const double xs[3] = {(double)tx0, (double)((tx0 + xe) / 2), (double)xe};
                const double ys[3] = {(double)ty0, (double)((ty0 + ye) / 2), (double)ye};
                h->Rmin = std::min(h->Rmin, c.R);
                h->Rmax = std::max(h->Rmax, c.R);
                for (double fx : xs) {
                    for (double fy : ys) {
                        const Frame f = pixel_frame(g, fx, fy);
                        const double d0 = f.u[0] - c.u[0], d1 = f.u[1] - c.u[1], d2 = f.u[2] - c.u[2];
                        du = std::max(du, std::sqrt(d0 * d0 + d1 * d1 + d2 * d2));
                        dR = std::max(dR, std::fabs(f.R - c.R));
                        h->Rmin = std::min(h->Rmin, f.R);
                        h->Rmax = std::max(h->Rmax, f.R);
                    }
                }
            }
        }
        // corners and edge midpoints sample the tile; u is smooth, margin 5%
        h->du_max[v] = 1.05 * du;
        h->dR_max[v] = 1.05 * dR;
    }
    // |r_p| is smallest at the point of the detector plane nearest the origin: the sampled minimum
    // above misses it when it lies inside a tile (a rotated or offset detector), so take it exactly
    // over the pixel range that any tile, quad or block reaches (one pixel before the first, up to
    // the end of the last 32 x 32 tile; the rotation keeps norms).
    {
        const double x0 = (-1.0 - g.half0) * g.pix + g.off0;
        const double x1 = (32.0 * ((g.n0 + 31) / 32) - g.half0) * g.pix + g.off0;
        const double y0 = (-1.0 - g.half1) * g.pix + g.off1;
        const double y1 = (32.0 * ((g.n1 + 31) / 32) - g.half1) * g.pix + g.off1;
        const double nx = std::min(std::max(0.0, x0), x1), ny = std::min(std::max(0.0, y0), y1);
        h->Rmin = std::min(h->Rmin, std::sqrt(nx * nx + ny * ny + g.detz * g.detz));
    }
    // Mirror-pair remainder: for a 2 x 2 quad the mean dq of a pair differs from the
    // quad mean by 1/4 of the mixed second difference, krev pix^2 |d_a d_b rh| / 4
    // (plus fourth-order terms).  With orthonormal detector axes a, b and
    // alpha = rh . a, beta = rh . b:  d_a d_b rh = (3 alpha beta rh - alpha b - beta a) / R^2,
    // so |d_a d_b rh| <= (sqrt(2) s + 3 s^2) / R^2 with s >= |alpha|, |beta|.  In the
    // detector frame alpha = dx / R and beta = dy / R (dx, dy the in-plane pixel
    // coordinates), largest in magnitude at the corners; one pixel is added for the
    // quads that reach past the last row or column.
    double emax = 0.0;
    for (double fx : {-1.0, (double)g.n0}) {
        for (double fy : {-1.0, (double)g.n1}) {
            emax = std::max(emax, std::fabs((fx - g.half0) * g.pix + g.off0));
            emax = std::max(emax, std::fabs((fy - g.half1) * g.pix + g.off1));
        }
    }
    const double sm = std::min(1.0, emax / h->Rmin);
    h->pair_coef = 1.05 * g.krev * g.pix * g.pix * (std::sqrt(2.0) * sm + 3.0 * sm * sm) /
                   (4.0 * h->Rmin * h->Rmin);
    // 2 x 4 block pairs: m_k - m_1 = (D2f[o_k] - D2f[o_1]) / 2 + O(o^4), where
    // |D2f[o]| <= krev pix^2 |o|^2 (1 + (2 + sqrt 2) s + 6 s^2) / R^2 from
    // d_a d_a rh = (3 alpha^2 rh - rh - 2 alpha a) / R^2 and the mixed derivative
    // above; |o_k|^2 <= 5/2, |o_1|^2 = 1/2.  Blocks reach three columns past the
    // last one.
    double ebl = emax;
    for (double fy : {-1.0, (double)g.n1 + 2.0})
        ebl = std::max(ebl, std::fabs((fy - g.half1) * g.pix + g.off1));
    const double sb = std::min(1.0, ebl / h->Rmin);
    h->block_coef = 1.05 * 1.5 * (1.0 + (2.0 + std::sqrt(2.0)) * sb + 6.0 * sb * sb) * g.krev * g.pix * g.pix /
                    (h->Rmin * h->Rmin);
    // Wave references (variants 9..12): d_k(p) - D_k = int_0^1 int_-1^1 D2f[a_k, o] dt ds along the
    // segment from the block centre P to the quad centre p, o = p - P, |a_k| = sqrt(2) / 2 pixel.
    // D2f is a symmetric bilinear form with |D2f[v, v]| <= krev pix^2 |v|^2 C(s) / R^2,
    // C(s) = 1 + (2 + sqrt 2) s + 6 s^2 (as above), so |D2f[a, o]| <= krev pix^2 |a| |o| C(s) / R^2.
    // In a 16 x 16 pixel wave block |o| <= 7 sqrt 2 pixels.  The segments and the block centres
    // reach to the edge of the last (padded) 32 x 32 tile, so s and R are taken over the padded
    // detector, R at its point nearest the foot of the detector normal.
    {
        const double x0 = (-1.0 - g.half0) * g.pix + g.off0;
        const double x1 = (32.0 * ((g.n0 + 31) / 32) - g.half0) * g.pix + g.off0;
        const double y0 = (-1.0 - g.half1) * g.pix + g.off1;
        const double y1 = (32.0 * ((g.n1 + 31) / 32) - g.half1) * g.pix + g.off1;
        const double ex = std::max(std::fabs(x0), std::fabs(x1)), ey = std::max(std::fabs(y0), std::fabs(y1));
        const double nx = std::min(std::max(0.0, x0), x1), ny = std::min(std::max(0.0, y0), y1);
        const double Rp = std::min(h->Rmin, std::sqrt(nx * nx + ny * ny + g.detz * g.detz));
        const double sr = std::min(1.0, std::max(ex, ey) / Rp);
        h->ref_coef = 1.05 * (1.0 + (2.0 + std::sqrt(2.0)) * sr + 6.0 * sr * sr) * g.krev * g.pix * g.pix *
                      (0.5 * std::sqrt(2.0)) * (7.0 * std::sqrt(2.0)) / (Rp * Rp);
        // 8 x 8 pixel sub-blocks (variant 14): the quads lie at |o| <= 3 sqrt 2 pixels from the
        // sub-block's reference quad, inside the same padded detector.  Replaced below by the
        // exact table where the detector is small enough.
        h->refs_coef = h->ref_coef * (3.0 / 7.0);
    }
    // Linear wave reference (variant 13): the kernel rotates by (ox Jx_k + oy Jy_k) . r' in place
    // of (d_k - D_k) . r'.  The residual d_k - D_k - ox Jx_k - oy Jy_k is evaluated here for every
    // quad that holds a detector pixel, with the FP32 D, Jx, Jy of the kernel (the same
    // wave_linear_ref), and its largest norm (with 5% margin for the FP64 rounding of this
    // evaluation) enters the bound.  Detectors beyond 2048 x 2048 pixels skip the table and the
    // variant.
    // The sub-block references of variants 14 and 15 are tabulated in the same pass: the largest
    // norm of d_k - D_k over every quad, D_k the FP32 half differences of the sub-block's reference
    // quad (lower corner 3 pixels into the 8 x 8 sub-block) as the kernel forms them, and for 15
    // the largest norm of the quad's m_0 - m_1 against the reference quad's (FP32 2 pi (m_0 - m_1)
    // as the kernel forms it).  Above 2048 x 2048 pixels variant 15 is not offered.
    h->lin_coef = std::numeric_limits<double>::infinity();
    h->cent_coef = std::numeric_limits<double>::infinity();
    if ((int64_t)g.n0 * g.n1 <= (int64_t)2048 * 2048) {
        double worst = 0.0, worst_sub = 0.0, worst_cent = 0.0;
        for (int tx0 = 0; tx0 < g.n0; tx0 += 32) {
            for (int ty0 = 0; ty0 < g.n1; ty0 += 32) {
                const int xe = std::min(tx0 + 32, g.n0) - 1, ye = std::min(ty0 + 32, g.n1) - 1;
                const Frame c = pixel_frame(g, 0.5 * (double)(tx0 + xe), 0.5 * (double)(ty0 + ye));
                for (int w = 0; w < 4; ++w) {
                    const int bx = tx0 + 16 * (w >> 1), by = ty0 + 16 * (w & 1);
                    if (bx >= g.n0 || by >= g.n1) continue;
                    float D[2][3], Jx[2][3], Jy[2][3];
                    wave_linear_ref(g, c.u, bx + 7.5, by + 7.5, D, Jx, Jy);
                    float Ds[4][2][3];
                    double Ms[4][3];
                    for (int sb = 0; sb < 4; ++sb) {
                        double hs[2][3], ms[3];
                        quad_half_diff(g, c.u, (double)(bx + 8 * (sb >> 1) + 3), (double)(by + 8 * (sb & 1) + 3), hs,
                                       ms);
                        for (int k = 0; k < 2; ++k)
                            for (int a = 0; a < 3; ++a) Ds[sb][k][a] = (float)hs[k][a];
                        for (int a = 0; a < 3; ++a) Ms[sb][a] = (double)(float)(kTwoPi * ms[a]) * kInv2Pi;
                    }
                    for (int i = 0; i < 8 && bx + 2 * i < g.n0; ++i) {
                        for (int j = 0; j < 8 && by + 2 * j < g.n1; ++j) {
                            double hq[2][3], mq[3];
                            quad_half_diff(g, c.u, (double)(bx + 2 * i), (double)(by + 2 * j), hq, mq);
                            const double ox = 2 * i - 7, oy = 2 * j - 7;
                            const int sb = ((i >> 2) << 1) | (j >> 2);
                            for (int k = 0; k < 2; ++k) {
                                double n2 = 0.0, s2 = 0.0;
                                for (int a = 0; a < 3; ++a) {
                                    const double r = hq[k][a] - (double)D[k][a] - ox * (double)Jx[k][a] -
                                                     oy * (double)Jy[k][a];
                                    const double rs = hq[k][a] - (double)Ds[sb][k][a];
                                    n2 += r * r;
                                    s2 += rs * rs;
                                }
                                worst = std::max(worst, std::sqrt(n2));
                                worst_sub = std::max(worst_sub, std::sqrt(s2));
                            }
                            double c2 = 0.0;
                            for (int a = 0; a < 3; ++a) c2 += (mq[a] - Ms[sb][a]) * (mq[a] - Ms[sb][a]);
                            worst_cent = std::max(worst_cent, std::sqrt(c2));
                        }
                    }
                }
            }
        }
        h->lin_coef = 1.05 * worst;
        h->refs_coef = 1.05 * worst_sub;
        h->cent_coef = 1.05 * worst_cent;
    }
}

// Error (revolutions) of folding W_cj into A_cj, i.e. of replacing lam_p W_cj
// by W_cj: W_max max|lam_p - 1|.
inline double fold_error(const Planner* h, int v, double rho)
{
    const Geom& g = h->g;
    if (!g.nf && !g.secondorder) return 0.0;
    const double dR = h->dR_max[v];
    const double Rlo = h->Rmin - dR - rho;
    if (!(Rlo > 0.0)) return std::numeric_limits<double>::infinity();
    const double W = g.krev * rho * rho / (2.0 * Rlo);
    return W * dR / (h->Rmin - dR);
}

// Bound of sigma_cj = (s - w) / D (near field; 0 otherwise) for tile variant v and emitters within
// |r_j| <= rho; infinite where the tile reaches the emitters.
inline double sigma_bound(const Planner* h, int v, double rho)
{
    const double Rlo = h->Rmin - h->dR_max[v] - rho;
    if (!(Rlo > 0.0)) return std::numeric_limits<double>::infinity();
    if (!h->g.nf) return 0.0;
    return (rho + rho * rho / (2.0 * Rlo)) / Rlo;
}

// Bound of the angle (rad) by which a lane rotates the reference phasor of variants 9..14 to its
// own pair: t = 2 pi (d_k - D_k) . r' with |r'| <= rho (1 + sigma); for the linear reference (13)
// the residual of its linear form is added.  0 for the tiles without a reference.
inline double ref_angle_bound(const Planner* h, int v, double rho)
{
    const int kind = tile_ref_kind(v);
    if (kind == 0) return 0.0;
    const double coef = kind >= 4 ? h->refs_coef : h->ref_coef + (kind == 3 ? h->lin_coef : 0.0);
    return kTwoPi * coef * rho * (1.0 + sigma_bound(h, v, rho));
}

// The truncation centre T^2 / 4 of variants 14 and 15, T the angle bound (0 for the others).  A lane
// whose own bound T_k (|t| <= T_k <= T over its emitters) exceeds T / sqrt 2 scales its sums by
// s = 1 / (1 + c), c = T_k^2 / 2 - T^2 / 4 in (0, T^2 / 4]; then s (1 + i t) - e^{i t} has real
// part t^2 / 2 - c + O(T^4), within [-T^2 / 4, T^2 / 4], and imaginary part at most 5 T^3 / 12.
// The other lanes keep s = 1 and their t^2 / 2 <= T^2 / 4.  The truncation is T^2 / 4 + T^3 / 2
// instead of T^2 / 2, and only the lanes far from their reference carry the scale's bias.
inline double ref_trunc_centre(const Planner* h, int v, double rho)
{
    if (tile_ref_kind(v) < 4) return 0.0;
    const double T = ref_angle_bound(h, v, rho);
    return std::isfinite(T) ? 0.25 * T * T : 0.0;
}

// A-priori bound of the per-term phase error (revolutions) of tile variant v
// for emitters within |r_j| <= rho.
inline double plan_bound(const Planner* h, int v, double rho)
{
    const Geom& g = h->g;
    const double du = h->du_max[v], dR = h->dR_max[v];
    const double Rlo = h->Rmin - dR - rho;
    if (!(Rlo > 0.0)) return std::numeric_limits<double>::infinity();
    const double Rhi = h->Rmax + dR + rho;
    const double ds = du * rho;
    double wmax = 0.0;
    const double sig = sigma_bound(h, v, rho);
    if (g.nf || g.secondorder) wmax = rho * rho / (2.0 * Rlo);
    const double W = g.krev * wmax;
    const double B = g.krev * ds * (1.0 + sig);
    const double tau = 0.5 + W + B;
    const double ulp = std::ldexp(1.0, std::ilogb(tau) - 23);
    const double u24 = std::ldexp(1.0, -24);
    double e = 2.0 * ulp;                // four FMA roundings of tau, each <= ulp/2
    e += 2.0 * u24 * B;                  // dq and r' rounded to FP32: sum_i |dq_i r_i| <= B
    e += 0.5 * u24 + 6.0 * u24 * W;      // A rounded; W formed in FP32; lam rounded
    e += 4.0 * u24 * B;                  // r' = r (1 + sigma) formed in FP32
    if (g.nf) {
        const double R3 = Rlo * Rlo * Rlo;
        const double wss = Rhi * Rhi / R3;
        const double wsR = (Rhi * rho + rho * rho) / R3;
        const double wRR = 2.0 * wmax * Rhi / R3;
        const double lin = 0.5 * wss * ds * ds + wsR * ds * dR + 0.5 * wRR * dR * dR +
                           wmax * dR * (dR + rho) / (Rlo * Rlo);
        e += g.krev * lin;
        if (g.scale) e += (ds / Rlo + rho * dR / (Rlo * Rlo)) * kInv2Pi;  // amplitude, as phase
    }
    e += kSinCosErrRev;
    const double fe = fold_error(h, v, rho);
    if (fe <= 0.1 * h->tol_rev) {
        e += fe;  // W folded into A
    } else if (tile_pair(v) && (g.nf || g.secondorder)) {
        return std::numeric_limits<double>::infinity();  // mirror pairs need the fold
    }
    if (tile_pair(v)) {
        const double rem = h->pair_coef * rho * (1.0 + sig);  // |e_k| = |(m_k - mc) . r'|
        if (v == kTileBlock) {
            // the three rotated centres, first order: |Delta| <= 2 pi block_coef rho (1 + sigma)
            const double delta = kTwoPi * h->block_coef * rho * (1.0 + sig);
            e += (0.5 * delta * delta + 4.0 * u24 * delta) * kInv2Pi;
        } else if (tile_pair_centred(v)) {
            // pair 1's centre phasor to first order: |Delta| = 2 pi |(m_0 - m_1) . r'| <= 4 pi rem,
            // error Delta^2 / 2 (rad), plus Delta's own FP32 rounding
            const double delta = 2.0 * kTwoPi * rem;
            e += (0.5 * delta * delta + 4.0 * u24 * delta) * kInv2Pi;
        } else {
            e += rem;  // dropped remainder e_k
        }
        e += kSinCosErrRev;  // second sincos (delta_k, or the reference E_k)
        e += 2.0 * u24;      // e^{i tau_c} (cos, sin) delta_k products, as phase
        if (const int ord = tile_ref_order(v)) {
            // rotation from the reference, t = 2 pi (d_k - D_k) . r': truncation t^2 / 2 (first
            // order), |t|^3 / 6 (second order) or T^2 / 4 + T^3 / 2 (first order with the sums of the
            // far lanes scaled, variant 14, plus the rounding of that product), t's own FP32
            // rounding, and the rounding of D_k and of the rotated (cos, sin).  The linear
            // reference (variant 13) rotates by the linear part of d_k - D_k instead: its
            // residual enters as a phase error of its own, and its slopes are formed and combined
            // in FP32 (8 roundings).
            const int kind = tile_ref_kind(v);
            const bool lin = kind == 3;
            const double t = ref_angle_bound(h, v, rho);
            const double trunc = kind >= 4 ? 0.25 * t * t + 0.5 * t * t * t + u24
                                           : (ord == 1 ? 0.5 * t * t : t * t * t / 6.0);
            e += (trunc + (lin ? 8.0 : 4.0) * u24 * t) * kInv2Pi + 2.0 * u24 * B + 2.0 * u24;
            if (lin) e += h->lin_coef * rho * (1.0 + sig);
            // variant 15 rotates pair 1 by the sub-block reference's (m_0 - m_1) . r' in place of
            // the quad's own: the difference is a phase error of pair 1
            if (kind == 5) e += h->cent_coef * rho * (1.0 + sig);
        }
    }
    return e;
}

// Tile variants of the FP32-split kernels whose error bound meets the tolerance.
inline std::vector<int> eligible_tiles(const Planner* h, double rho, double* best_bound)
{
    std::vector<int> out;
    *best_bound = std::numeric_limits<double>::infinity();
    for (int v = 0; v < kNumTiles; ++v) {
        if (v == kTileNone || (tile_pair(v) && h->summode != 0)) continue;
        const double b = plan_bound(h, v, rho);
        *best_bound = std::min(*best_bound, b);
        if (b <= h->tol_rev) out.push_back(v);
    }
    return out;
}

}  // namespace idisim
