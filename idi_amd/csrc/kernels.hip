// Forward kernels of the direct-sum wavefield for gfx950 (wave64, 256 CUs).
//
// FP32-split kernels (reference counterpart: wfkernelf, sim.cu:101-144)
//   - a workgroup (256 lanes = 4 waves) owns a tile of pixels; each lane owns
//     PX x PY neighbouring pixels;
//   - emitters stream through LDS in chunks of 256: each lane forms ONE
//     emitter's record in FP64 (range-reduced linear phase A, curvature W,
//     scaled position r', amplitude), all lanes then read every record as an
//     LDS broadcast;
//   - per pair: tau = A + lam W + dq . r' (4 FMA), v_sin_f32 / v_cos_f32 on
//     revolutions, 2 FMA into per-chunk FP32 partial sums, which are merged
//     into the running sum with Kahan compensation after every chunk (or, with
//     SUM == 1, Kahan on every term as cpu.py:97-105 does);
//   - emitters can be split over workgroups (blockIdx.y); partial sums then go
//     to a scratch buffer and a finalize kernel combines them in FP64.
// FP64 kernel (reference counterpart: wfkerneld, sim.cu:146-149, and the numba
// kernel cpu.py:68-107): the literal formula in FP64 with Kahan per term, same
// tiling and LDS staging; it is the "plain FP64" comparator and the fallback.
#include <type_traits>

#include "device_util.hpp"
#include "sincos64.hpp"

namespace idisim {

// ---------------------------------------------------------------------------
// FP32-split tile: accumulate emitters [e0, e1) for this lane's pixels.
template <int PX, int PY, bool HAS_W, bool HAS_AMP, int SUM>
__device__ __forceinline__ void split_tile(const Geom& g, const double* __restrict__ emit,
                                           const double* __restrict__ amp, int64_t e0, int64_t e1,
                                           int tx0, int ty0, float4* s0, float2* s1,
                                           float (&acc_re)[PX * PY], float (&acc_im)[PX * PY],
                                           float (&cmp_re)[PX * PY], float (&cmp_im)[PX * PY])
{
    constexpr int P = PX * PY;
    constexpr int TX = 16 * PX, TY = 16 * PY;
    const int tid = threadIdx.x;
    const int lx = tid >> 4, ly = tid & 15;

    double ccx, ccy;
    tile_centre(g, tx0, ty0, TX, TY, ccx, ccy);
    const Frame c = pixel_frame(g, ccx, ccy);

    float dqx[P], dqy[P], dqz[P], lam[P];
#pragma unroll
    for (int i = 0; i < PX; ++i) {
#pragma unroll
        for (int j = 0; j < PY; ++j) {
            const int p = i * PY + j;
            const Frame f = pixel_frame(g, (double)(tx0 + lx * PX + i), (double)(ty0 + ly * PY + j));
            dqx[p] = (float)(-g.krev * (f.u[0] - c.u[0]));
            dqy[p] = (float)(-g.krev * (f.u[1] - c.u[1]));
            dqz[p] = (float)(-g.krev * (f.u[2] - c.u[2]));
            lam[p] = (float)(c.R / f.R);
            acc_re[p] = 0.f;
            acc_im[p] = 0.f;
            cmp_re[p] = 0.f;
            cmp_im[p] = 0.f;
        }
    }

    for (int64_t base = e0; base < e1; base += kChunk) {
        const int cnt = (e1 - base) < (int64_t)kChunk ? (int)(e1 - base) : kChunk;
        __syncthreads();  // the previous chunk has been consumed by every lane
        if (tid < cnt) {
            const int64_t e = base + tid;
            const double* a4 = emit + 4 * e;
            const double ea = (HAS_AMP && amp != nullptr) ? amp[e] : 1.0;
            const SplitRec r = make_split_rec(g, c, a4[0], a4[1], a4[2], a4[3], ea);
            s0[tid] = make_float4(r.x, r.y, r.z, r.A);
            s1[tid] = make_float2(r.W, r.amp);
        }
        __syncthreads();

        float sre[P], sim[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            sre[p] = 0.f;
            sim[p] = 0.f;
        }
#pragma unroll 2
        for (int e = 0; e < cnt; ++e) {
            const float4 r0 = s0[e];
            float2 r1 = make_float2(0.f, 1.f);
            if (HAS_W || HAS_AMP) r1 = s1[e];
#pragma unroll
            for (int p = 0; p < P; ++p) {
                float t = HAS_W ? fmaf(lam[p], r1.x, r0.w) : r0.w;
                t = fmaf(dqx[p], r0.x, t);
                t = fmaf(dqy[p], r0.y, t);
                t = fmaf(dqz[p], r0.z, t);
                const float sn = __builtin_amdgcn_sinf(t);
                const float cs = __builtin_amdgcn_cosf(t);
                if (SUM == 0) {
                    if (HAS_AMP) {
                        sre[p] = fmaf(r1.y, cs, sre[p]);
                        sim[p] = fmaf(r1.y, sn, sim[p]);
                    } else {
                        sre[p] += cs;
                        sim[p] += sn;
                    }
                } else {
                    kahan_add(acc_re[p], cmp_re[p], HAS_AMP ? r1.y * cs : cs);
                    kahan_add(acc_im[p], cmp_im[p], HAS_AMP ? r1.y * sn : sn);
                }
            }
        }
        if (SUM == 0) {
#pragma unroll
            for (int p = 0; p < P; ++p) {
                kahan_add(acc_re[p], cmp_re[p], sre[p]);
                kahan_add(acc_im[p], cmp_im[p], sim[p]);
            }
        }
    }
}

template <int PX, int PY, bool HAS_W, bool HAS_AMP, int SUM>
__global__ __launch_bounds__(kBlock) void k_fwd_split(Geom g, const double* __restrict__ emit,
                                                      const double* __restrict__ amp, int64_t n,
                                                      int tiles_y, int64_t per_split,
                                                      float4* __restrict__ partial, void* out, int kind)
{
    constexpr int P = PX * PY;
    constexpr int TX = 16 * PX, TY = 16 * PY;
    __shared__ float4 s0[kChunk];
    __shared__ float2 s1[kChunk];
    const int tile = blockIdx.x;
    const int split = blockIdx.y;
    const int tx0 = (tile / tiles_y) * TX, ty0 = (tile % tiles_y) * TY;
    const int64_t e0 = (int64_t)split * per_split;
    const int64_t e1 = (e0 + per_split < n) ? e0 + per_split : n;
    float re[P], im[P], cre[P], cim[P];
    split_tile<PX, PY, HAS_W, HAS_AMP, SUM>(g, emit, amp, e0, e1, tx0, ty0, s0, s1, re, im, cre, cim);

    const int lx = threadIdx.x >> 4, ly = threadIdx.x & 15;
    const size_t npix = (size_t)g.n0 * g.n1;
#pragma unroll
    for (int i = 0; i < PX; ++i) {
#pragma unroll
        for (int j = 0; j < PY; ++j) {
            const int p = i * PY + j;
            const int x = tx0 + lx * PX + i, y = ty0 + ly * PY + j;
            if (x < g.n0 && y < g.n1) {
                const size_t idx = (size_t)x * g.n1 + y;
                if (partial != nullptr)
                    partial[(size_t)split * npix + idx] = make_float4(re[p], im[p], cre[p], cim[p]);
                else
                    store_split(g, x, y, (double)re[p] - (double)cre[p], (double)im[p] - (double)cim[p],
                                out, kind, idx);
            }
        }
    }
}

// Persistent batched kernel: workgroups pull (shot, tile) items from a queue.
// Items are sorted by emitter count (largest first) on the host, so the tail
// of the queue is made of short items.  Every workgroup leaves the loop once
// the counter passes nitems.

template <int PX, int PY, bool HAS_W, bool HAS_AMP, int SUM>
__global__ __launch_bounds__(kBlock) void k_batched_split(Geom g0, const double* __restrict__ emit,
                                                          const double* __restrict__ amp,
                                                          const WorkItem* __restrict__ items, int nitems,
                                                          int* __restrict__ counter, int tiles_y, void* out,
                                                          int kind)
{
    constexpr int P = PX * PY;
    constexpr int TX = 16 * PX, TY = 16 * PY;
    __shared__ float4 s0[kChunk];
    __shared__ float2 s1[kChunk];
    __shared__ int s_item;
    const int lx = threadIdx.x >> 4, ly = threadIdx.x & 15;
    const size_t npix = (size_t)g0.n0 * g0.n1;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = s_item;
        if (it >= nitems) break;
        const WorkItem w = items[it];
        const Geom g = item_geom();
        const int tx0 = (w.tile / tiles_y) * TX, ty0 = (w.tile % tiles_y) * TY;
        float re[P], im[P], cre[P], cim[P];
        split_tile<PX, PY, HAS_W, HAS_AMP, SUM>(g, emit, amp, w.e0, w.e1, tx0, ty0, s0, s1, re, im, cre,
                                                cim);
        char* outs = static_cast<char*>(out) +
                     (size_t)w.shot * npix * (kind == OUT_INTENSITY ? sizeof(float) : sizeof(float2));
        const Geom gs = item_geom();  // store frames formed here (see k_batched_pair)
#pragma unroll
        for (int i = 0; i < PX; ++i) {
#pragma unroll
            for (int j = 0; j < PY; ++j) {
                const int p = i * PY + j;
                const int x = tx0 + lx * PX + i, y = ty0 + ly * PY + j;
                if (x < gs.n0 && y < gs.n1)
                    store_split(gs, x, y, (double)re[p] - (double)cre[p], (double)im[p] - (double)cim[p],
                                outs, kind, (size_t)x * gs.n1 + y);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// FP32-split tile with mirror pairs (tile variants 5 and 6).  Each lane owns QY
// quads of 2 x 2 pixels (a 2 x 2QY block; the 16 x 16 lanes cover 32 x 32QY
// pixels).  In a quad, pixels (0,0)-(1,1) (pair 0) and (1,0)-(0,1) (pair 1) are
// mirror images about the quad centre.  With dq the tile-relative phase gradient
// of a pixel, mc the mean dq of the quad and d_k half the difference of pair k's
// dq (second member minus first):
//   tau = tau_c + e_k -/+ delta_k,  tau_c = A + mc . r',  delta_k = d_k . r',
// where e_k = (mean dq of pair k - mc) . r' is a second-order remainder.  For a
// square quad the isotropic curvature cancels in e_k and what is left is +/- 1/4
// of the mixed second difference of dq, which the host planner bounds and which
// is left out.  Per quad and emitter: one sincos of tau_c; per pair one sincos of
// delta_k and four FMA into
//   U_k = sum e^{i tau_c} cos delta_k,   V_k = sum e^{i tau_c} sin delta_k,
// and the pair's fields are U_k - i V_k (first member) and U_k + i V_k (second).
// That is 6 transcendentals and 17 FMA per 4 pixels where split_tile spends 8
// and 20.  The W term must be folded into A (no lam_p per pixel); sums are the
// chunked Kahan sums of split_tile.
// CENTRED: each pair gets its own centre, the pair's mean dq m_k, so that no
// remainder is dropped.  tau_c is then pair 0's centre, and pair 1's phasor is
// e^{i tau_c} e^{-i Delta} with Delta = 2 pi (m_0 - m_1) . r' (a few 1e-4 rad),
// taken to first order as (1 - i Delta): 5 more FMA per quad, an error of
// Delta^2 / 2 that the host planner bounds.
// REF (1 or 2): the sincos of delta_k comes from a reference.  Every wave owns a
// 16 x 16QY pixel block (8 x 8 lanes) and D_k, the half difference d_k evaluated at
// the block centre (pixels at +/- (1/2, 1/2) about it); the record stage forms
// E_k = e^{i 2 pi D_k . r'} once per emitter and wave (8 sincos per emitter, 1/128
// per pixel pair), and each lane reaches its own pair by the rotation
// e^{i t}, t = 2 pi (d_k - D_k) . r' (|t| a few 1e-3 rad), taken to first order
// (1 + i t; REF 1) or second order (1 + i t - t^2 / 2; REF 2).  Per quad and
// emitter that is 2 transcendentals and 21 (REF 1) or 27 (REF 2) other VALU where
// REF 0 spends 6 and 17 (plus 2 multiplies by the amplitude, which with a reference
// is folded into E_k instead).  The host planner bounds t^2 / 2 (REF 1) or |t|^3 / 6
// (REF 2).  sD holds each wave's D_k (FP32, 6 floats of 18 per wave; REF 3 adds its slopes)
// and sE their phasors, (cos, sin) of both pairs per float4, with REF 2 a second float4
// scaled by sqrt 2.
// Emitters staged in LDS per round by the mirror-pair tiles (records and, with a wave reference,
// the four waves' phasors: 16 B x (1 + 4 or 8) per emitter).
constexpr int kPairChunk = 256;
// REF 4 (variant 14) takes a reference for each 8 x 8 pixel sub-block of a wave (four per wave)
// and rotates to first order; the lanes far from their reference scale their sums at the end
// (g.refrho, g.refc).  Its 16 phasors per emitter are stored emitter-major with a stride of
// 17 float4 (a wave's four sub-blocks read adjacent float4s; the record stage's writes from
// consecutive lanes fall in distinct banks), 128 emitters per round, so that the LDS of a
// workgroup stays under 40 KB (4 per CU).
constexpr int kSubRefs = 16;
constexpr int kSubStride = 17;
// REF 3 (variant 13) holds two float4 per emitter and wave: 256 emitters per round take 36 KB of
// LDS, four workgroups per CU.  224 (five per CU under 32 KB) measured the same on 100k emitters,
// but a 256-emitter split then takes two rounds, the second of 32 emitters (4096 emitters over
// 16 splits: 5 us of a 51 us launch).
constexpr int kRef3Chunk = 256;
// REF 5 (variant 15) adds one float per emitter and sub-block, the rotation between the pair
// centres, 2 pi (m_0 - m_1) . r' of the sub-block's reference quad: 112 emitters per round.
constexpr int kRef5Chunk = 112;
__host__ __device__ constexpr int pair_chunk(int ref)
{
    return ref == 4 ? 128 : (ref == 5 ? kRef5Chunk : (ref == 3 ? kRef3Chunk : kPairChunk));
}
// per sub-block: D_0, D_1 (REF 4), and 2 pi (m_0 - m_1) (REF 5)
__host__ __device__ constexpr int pair_sd_floats(int ref)
{
    return ref == 5 ? 9 * kSubRefs : (ref == 4 ? 6 * kSubRefs : (ref ? 72 : 1));
}
__host__ __device__ constexpr int pair_se_float4(int ref)
{
    return ref == 0 ? 1 : (ref == 1 ? 4 * kPairChunk : (ref >= 4 ? kSubStride * pair_chunk(ref) : 8 * pair_chunk(ref)));
}
__host__ __device__ constexpr int pair_sc_floats(int ref) { return ref == 5 ? kSubRefs * kRef5Chunk : 1; }
// (W, amplitude) of each staged emitter: read only by the tiles without a reference (the
// others fold the curvature into A and the amplitude into their reference phasors)
__host__ __device__ constexpr int pair_s1_float2(int ref) { return ref == 0 ? kPairChunk : 1; }

template <int REF>
__device__ __forceinline__ void pair_lane(int tid, int& lx, int& ly)
{
    if (REF == 0) {
        lx = tid >> 4;
        ly = tid & 15;
    } else {
        const int w = tid >> 6, l = tid & 63;
        lx = ((w >> 1) << 3) | (l >> 3);
        ly = ((w & 1) << 3) | (l & 7);
    }
}

// Set-up of a mirror-pair tile from the detector geometry alone: the tile-centre frame c, each
// lane's quad terms mc, dv and d (d relative to the lane's reference), and the workgroup's
// reference data in sD (REF >= 1).  REF 3 stages its half differences in sE and synchronises the
// workgroup.  About nine FP64 pixel frames per lane: k_pair_table forms them once per handle and
// tile variant, and k_fwd_pair then loads them (pair_tab_floats).
template <int QY, bool CENTRED, int REF>
__device__ __forceinline__ void pair_setup(const Geom& g, int tx0, int ty0, int tid, int lx, int ly, int myref,
                                           float* sD, float4* sE, Frame& c, float (&mc)[QY][3],
                                           float (&dv)[QY][3], float (&d)[QY][2][3])
{
    constexpr int TX = 32, TY = 32 * QY;
    double ccx, ccy;
    tile_centre(g, tx0, ty0, TX, TY, ccx, ccy);
    c = pixel_frame(g, ccx, ccy);

    // the wave's reference half differences D_k (as the FP32 values the records use)
    double Dref[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    // REF 3: the wave's linear reference (D_k and its slopes, 18 floats per wave) formed by one
    // lane of the wave
    if (REF == 3) {
        static_assert(REF != 3 || QY == 1, "the linear wave reference has one quad per lane");
        const int w = tid >> 6, l = tid & 63;
        // lanes 0..4 of each wave form the half differences of the quads centred at the block
        // centre and one pixel to either side (the terms of wave_linear_ref), staged in the
        // phasor area of LDS, which no lane reads before the first chunk's records
        double* sH = reinterpret_cast<double*>(sE);
        if (l < 5) {
            const double px = (double)(tx0 + 16 * (w >> 1)) + 7.5;
            const double py = (double)(ty0 + 16 * (w & 1)) + 7.5;
            const double sx = l == 1 ? 0.5 : (l == 2 ? -1.5 : -0.5);
            const double sy = l == 3 ? 0.5 : (l == 4 ? -1.5 : -0.5);
            double h[2][3];
            quad_half_diff(g, c.u, px + sx, py + sy, h);
            for (int k = 0; k < 2; ++k)
                for (int a = 0; a < 3; ++a) sH[(w * 5 + l) * 6 + k * 3 + a] = h[k][a];
        }
        __syncthreads();
        if (l == 0) {  // the combination of wave_linear_ref, as FP32 values
            const double* H = sH + w * 30;
            for (int i = 0; i < 6; ++i) {
                sD[w * 18 + i] = (float)H[i];
                sD[w * 18 + 6 + i] = (float)(0.5 * (H[6 + i] - H[12 + i]));
                sD[w * 18 + 12 + i] = (float)(0.5 * (H[18 + i] - H[24 + i]));
            }
        }
    }
    constexpr bool SUB = REF == 4 || REF == 5;  // a reference per 8 x 8 pixel sub-block
    constexpr int DSTR = REF == 5 ? 9 : 6;       // floats per sub-block in sD
    if (REF == 1 || REF == 2 || SUB) {
        static_assert(!SUB || (QY == 1 && CENTRED), "sub-block references: one centred quad per lane");
        const int w = tid >> 6, l = tid & 63;
        const int sb = myref & 3;
        // the reference quad's centre: the wave block's centre (REF 1, 2) or the sub-block's
        const double px = SUB ? (double)(tx0 + 16 * (w >> 1) + 8 * (sb >> 1)) + 3.5
                              : (double)(tx0 + 16 * (w >> 1)) + 7.5;
        const double py = SUB ? (double)(ty0 + 16 * (w & 1) + 8 * (sb & 1)) + 3.5
                              : (double)(ty0 + 16 * QY * (w & 1)) + (8.0 * QY - 0.5);
        // the lane that stores its reference: the first of the wave (REF 1, 2) or of the sub-block
        const bool writer = SUB ? (l & 27) == 0 : l == 0;
        float* Dst = SUB ? sD + DSTR * myref : sD + 18 * w;
        const Frame f00 = pixel_frame(g, px - 0.5, py - 0.5), f11 = pixel_frame(g, px + 0.5, py + 0.5);
        const Frame f10 = pixel_frame(g, px + 0.5, py - 0.5), f01 = pixel_frame(g, px - 0.5, py + 0.5);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double q00 = -g.krev * (f00.u[a] - c.u[a]), q11 = -g.krev * (f11.u[a] - c.u[a]);
            const double q10 = -g.krev * (f10.u[a] - c.u[a]), q01 = -g.krev * (f01.u[a] - c.u[a]);
            const float D0 = (float)(0.5 * (q11 - q00)), D1 = (float)(0.5 * (q01 - q10));
            Dref[0][a] = (double)D0;
            Dref[1][a] = (double)D1;
            if (writer) {
                Dst[a] = D0;
                Dst[3 + a] = D1;
                // REF 5: the reference quad's 2 pi (m_0 - m_1), as the lanes' dv below
                if (REF == 5) Dst[6 + a] = (float)(kTwoPi * (0.5 * (q00 + q11) - 0.5 * (q10 + q01)));
            }
        }
    }
    // REF 1, 4: t in radians; REF 2: t / sqrt 2, so that its square is t^2 / 2
    const double dscale = REF == 2 ? kTwoPi * 0.70710678118654752440 : kTwoPi;

#pragma unroll
    for (int q = 0; q < QY; ++q) {
        double dq[2][2][3];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const Frame f = pixel_frame(g, (double)(tx0 + 2 * lx + i), (double)(ty0 + 2 * QY * ly + 2 * q + j));
#pragma unroll
                for (int a = 0; a < 3; ++a) dq[i][j][a] = -g.krev * (f.u[a] - c.u[a]);
            }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double m0 = 0.5 * (dq[0][0][a] + dq[1][1][a]);
            const double m1 = 0.5 * (dq[1][0][a] + dq[0][1][a]);
            mc[q][a] = (float)(CENTRED ? m0 : 0.5 * (m0 + m1));
            dv[q][a] = (float)(kTwoPi * (m0 - m1));
            const double h0 = 0.5 * (dq[1][1][a] - dq[0][0][a]);
            const double h1 = 0.5 * (dq[0][1][a] - dq[1][0][a]);
            d[q][0][a] = (float)(REF ? dscale * (h0 - Dref[0][a]) : h0);
            d[q][1][a] = (float)(REF ? dscale * (h1 - Dref[1][a]) : h1);
        }
    }
}

// The lane's sub-block (REF 4, 5): 4 w + (2 x half + y half of the wave's 16 x 16 pixel block).
__device__ __forceinline__ int pair_subblock(int tid)
{
    const int w = tid >> 6, l = tid & 63;
    return 4 * w + (((l >> 5) << 1) | ((l >> 2) & 1));
}

// Table of pair_setup for one tile (one quad per lane), in floats: the lanes' mc, dv, d as three
// float4 planes of kBlock, the reference data sD (padded to whole float4), and the tile-centre
// frame as 8 doubles (R, rh, u, R - z).
constexpr int kPairTabLane = 3 * 4 * kBlock;
__host__ __device__ constexpr int pair_tab_sd(int ref) { return ref == 0 ? 0 : (pair_sd_floats(ref) + 3) / 4 * 4; }
__host__ __device__ constexpr int pair_tab_floats(int ref) { return kPairTabLane + pair_tab_sd(ref) + 16; }

// PRIO: the wave priority follows the share of the workgroup's emitters done (prio_by_progress).
// tab: this tile's pair_setup table (QY 1), or nullptr to form it here.
template <int QY, bool HAS_AMP, bool CENTRED, int REF, bool PRIO = false>
__device__ __forceinline__ void pair_tile(const Geom& g, const double* __restrict__ emit,
                                          const double* __restrict__ amp, int64_t e0, int64_t e1, int tx0,
                                          int ty0, const float* __restrict__ tab, float4* s0, float2* s1,
                                          float* sD, float4* sE, float* sC, float (&acc)[QY][2][4],
                                          float (&cmp)[QY][2][4])
{
    const int tid = threadIdx.x;
    int lx, ly;
    pair_lane<REF>(tid, lx, ly);
    if (PRIO) __builtin_amdgcn_s_setprio(3);

    // REF 3: the quad centre's offset from the wave block's centre (pixels)
    float ox = 0.f, oy = 0.f;
    if (REF == 3) {
        const int l = tid & 63;
        ox = (float)(2 * (l >> 3) - 7);
        oy = (float)(2 * (l & 7) - 7);
    }
    constexpr bool SUB = REF == 4 || REF == 5;  // a reference per 8 x 8 pixel sub-block
    const int myref = SUB ? pair_subblock(tid) : 0;

    Frame c;
    float mc[QY][3], d[QY][2][3], dv[QY][3];
    bool loaded = false;
    if constexpr (QY == 1) {
        if (tab != nullptr) {
            const float4* tl = reinterpret_cast<const float4*>(tab);
            const float4 t0 = tl[tid], t1 = tl[kBlock + tid], t2 = tl[2 * kBlock + tid];
            mc[0][0] = t0.x, mc[0][1] = t0.y, mc[0][2] = t0.z, dv[0][0] = t0.w;
            dv[0][1] = t1.x, dv[0][2] = t1.y, d[0][0][0] = t1.z, d[0][0][1] = t1.w;
            d[0][0][2] = t2.x, d[0][1][0] = t2.y, d[0][1][1] = t2.z, d[0][1][2] = t2.w;
            // sD is read after the first chunk's barrier
            for (int i = tid; i < pair_tab_sd(REF); i += kBlock)
                if (i < pair_sd_floats(REF)) sD[i] = tab[kPairTabLane + i];
            const double* fr = reinterpret_cast<const double*>(tab + kPairTabLane + pair_tab_sd(REF));
            c.R = fr[0];
            c.rh[0] = fr[1], c.rh[1] = fr[2], c.rh[2] = fr[3];
            c.u[0] = fr[4], c.u[1] = fr[5], c.u[2] = fr[6];
            c.RmZ = fr[7];
            loaded = true;
        }
    }
    if (!loaded) pair_setup<QY, CENTRED, REF>(g, tx0, ty0, tid, lx, ly, myref, sD, sE, c, mc, dv, d);
    constexpr int DSTR = REF == 5 ? 9 : 6;  // floats per sub-block in sD
#pragma unroll
    for (int q = 0; q < QY; ++q) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                acc[q][k][m] = 0.f;
                cmp[q][k][m] = 0.f;
            }
        }
    }
    constexpr int CH = pair_chunk(REF);
    // this lane's phasors: its wave's (REF 1..3, emitter stride 1) or its sub-block's (REF 4, 5)
    constexpr int ESTR = SUB ? kSubStride : 1;
    const float4* eref = SUB ? sE + myref : sE + ((REF == 1 ? 1 : 2) * (tid >> 6)) * CH;

    for (int64_t base = e0; base < e1; base += CH) {
        const int cnt = (e1 - base) < (int64_t)CH ? (int)(e1 - base) : CH;
        // REF 4, 5: lanes tid and tid + 128 both form the record of emitter tid and each stores 8
        // of its 16 sub-block phasors
        const int el = SUB ? (tid & 127) : tid;
        __syncthreads();  // the previous chunk has been consumed by every lane
        if (PRIO) __builtin_amdgcn_s_setprio(3);  // the records (FP64): every wave waits for them
        if (el < cnt) {
            const int64_t e = base + el;
            const double* a4 = emit + 4 * e;
            const double ea = (HAS_AMP && amp != nullptr) ? amp[e] : 1.0;
            const SplitRec r = make_split_rec(g, c, a4[0], a4[1], a4[2], a4[3], ea);
            if (!SUB || tid < 128) s0[el] = make_float4(r.x, r.y, r.z, r.A);
            if (REF == 0) s1[el] = make_float2(r.W, r.amp);
            if (SUB) {
                const float sc = HAS_AMP ? r.amp : 1.f;
                const int first = (tid >> 7) * (kSubRefs / 2);
#pragma unroll 2
                for (int j = 0; j < kSubRefs / 2; ++j) {
                    const float* D = sD + DSTR * (first + j);
                    float dl0 = D[0] * r.x, dl1 = D[3] * r.x;
                    dl0 = fmaf(D[1], r.y, dl0);
                    dl1 = fmaf(D[4], r.y, dl1);
                    dl0 = fmaf(D[2], r.z, dl0);
                    dl1 = fmaf(D[5], r.z, dl1);
                    sE[el * kSubStride + first + j] =
                        make_float4(sc * __builtin_amdgcn_cosf(dl0), sc * __builtin_amdgcn_sinf(dl0),
                                    sc * __builtin_amdgcn_cosf(dl1), sc * __builtin_amdgcn_sinf(dl1));
                    if (REF == 5) {
                        float dc = D[6] * r.x;
                        dc = fmaf(D[7], r.y, dc);
                        dc = fmaf(D[8], r.z, dc);
                        sC[el * kSubRefs + first + j] = dc;
                    }
                }
            } else if (REF == 3) {
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    const float* D = sD + w * 18;
                    float v[12];  // (cos, sin) of both pairs, then the slopes 2 pi Jx_k . r', 2 pi Jy_k . r'
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        float dl = D[k * 3] * r.x;
                        dl = fmaf(D[k * 3 + 1], r.y, dl);
                        dl = fmaf(D[k * 3 + 2], r.z, dl);
                        const float sn = __builtin_amdgcn_sinf(dl), cs = __builtin_amdgcn_cosf(dl);
                        v[2 * k] = HAS_AMP ? r.amp * cs : cs;
                        v[2 * k + 1] = HAS_AMP ? r.amp * sn : sn;
#pragma unroll
                        for (int s2 = 0; s2 < 2; ++s2) {
                            const float* J = D + 6 + 6 * s2 + k * 3;
                            float gl = J[0] * r.x;
                            gl = fmaf(J[1], r.y, gl);
                            gl = fmaf(J[2], r.z, gl);
                            v[4 + 2 * k + s2] = (float)kTwoPi * gl;
                        }
                    }
                    sE[(2 * w) * CH + tid] = make_float4(v[0], v[1], v[2], v[3]);
                    sE[(2 * w + 1) * CH + tid] = make_float4(v[4], v[5], v[6], v[7]);
                }
            } else if (REF != 0) {
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    float ec[2], es[2];
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        const float* D = sD + w * 18 + k * 3;
                        float dl = D[0] * r.x;
                        dl = fmaf(D[1], r.y, dl);
                        dl = fmaf(D[2], r.z, dl);
                        // the amplitude rides on the phasor: the lanes' products are linear in it
                        es[k] = HAS_AMP ? r.amp * __builtin_amdgcn_sinf(dl) : __builtin_amdgcn_sinf(dl);
                        ec[k] = HAS_AMP ? r.amp * __builtin_amdgcn_cosf(dl) : __builtin_amdgcn_cosf(dl);
                    }
                    sE[(REF * w) * CH + tid] = make_float4(ec[0], es[0], ec[1], es[1]);
                    if (REF == 2) {
                        constexpr float r2 = 1.41421356237309504880f;
                        sE[(REF * w + 1) * CH + tid] =
                            make_float4(r2 * ec[0], r2 * es[0], r2 * ec[1], r2 * es[1]);
                    }
                }
            }
        }
        __syncthreads();

        float su[QY][2][4];
#pragma unroll
        for (int q = 0; q < QY; ++q)
#pragma unroll
            for (int k = 0; k < 2; ++k)
#pragma unroll
                for (int m = 0; m < 4; ++m) su[q][k][m] = 0.f;
        // PRIO: blocks of 32 emitters, each at the priority of the share done before it
        constexpr int EB = 32;
        for (int eb = 0; eb < (PRIO ? cnt : 1); eb += EB) {  // one pass without PRIO
        if (PRIO) prio_by_progress((int)((32 * (base - e0 + eb)) / (e1 - e0)));
        const int ee = !PRIO ? cnt : (cnt - eb < EB ? cnt : eb + EB);
#pragma unroll (SUB ? 4 : 2)  // SUB: one address update per four emitters (the VGPR bound holds)
        for (int e = eb; e < ee; ++e) {
            const float4 r0 = s0[e];
            const float am = (HAS_AMP && REF == 0) ? s1[e].y : 1.f;
            float4 E = make_float4(0.f, 0.f, 0.f, 0.f), E2 = E;
            if (REF != 0) E = eref[e * ESTR];
            if (REF == 2 || REF == 3) E2 = eref[CH + e];  // REF 2: sqrt 2 E; REF 3: the slopes
#pragma unroll
            for (int q = 0; q < QY; ++q) {
                float t = fmaf(mc[q][0], r0.x, r0.w);
                t = fmaf(mc[q][1], r0.y, t);
                t = fmaf(mc[q][2], r0.z, t);
                float sc = __builtin_amdgcn_sinf(t);
                float cc = __builtin_amdgcn_cosf(t);
                if (HAS_AMP && REF == 0) {  // with a reference the amplitude is in E
                    sc *= am;
                    cc *= am;
                }
                float pc[2] = {cc, cc}, ps[2] = {sc, sc};  // centre phasor of each pair
                if (CENTRED) {
                    float dl;
                    if (REF == 5) {  // the sub-block reference's rotation between the pair centres
                        dl = sC[e * kSubRefs + myref];
                    } else {
                        dl = dv[q][0] * r0.x;
                        dl = fmaf(dv[q][1], r0.y, dl);
                        dl = fmaf(dv[q][2], r0.z, dl);
                    }
                    pc[1] = fmaf(dl, sc, cc);   // (cc + i sc)(1 - i dl)
                    ps[1] = fmaf(-dl, cc, sc);
                }
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    float dl;
                    if (REF == 3) {
                        dl = fmaf(oy, k ? E2.w : E2.y, ox * (k ? E2.z : E2.x));
                    } else {
                        dl = d[q][k][0] * r0.x;
                        dl = fmaf(d[q][k][1], r0.y, dl);
                        dl = fmaf(d[q][k][2], r0.z, dl);
                    }
                    float sd, cd;
                    if (REF == 0) {
                        sd = __builtin_amdgcn_sinf(dl);
                        cd = __builtin_amdgcn_cosf(dl);
                    } else {
                        const float ec = k ? E.z : E.x, es = k ? E.w : E.y;
                        if (REF != 2) {  // (ec + i es)(1 + i t), t = dl
                            cd = fmaf(-dl, es, ec);
                            sd = fmaf(dl, ec, es);
                        } else {         // (ec + i es)(1 - u^2 + i sqrt2 u), u = dl
                            const float sec = k ? E2.z : E2.x, ses = k ? E2.w : E2.y;
                            const float uu = dl * dl;
                            cd = fmaf(-dl, ses, fmaf(-uu, ec, ec));
                            sd = fmaf(dl, sec, fmaf(-uu, es, es));
                        }
                    }
                    su[q][k][0] = fmaf(pc[k], cd, su[q][k][0]);
                    su[q][k][1] = fmaf(ps[k], cd, su[q][k][1]);
                    su[q][k][2] = fmaf(pc[k], sd, su[q][k][2]);
                    su[q][k][3] = fmaf(ps[k], sd, su[q][k][3]);
                }
            }
        }
        }
#pragma unroll
        for (int q = 0; q < QY; ++q)
#pragma unroll
            for (int k = 0; k < 2; ++k)
#pragma unroll
                for (int m = 0; m < 4; ++m) kahan_add(acc[q][k][m], cmp[q][k][m], su[q][k][m]);
    }
    if (SUB) {
        // the lane's own angle bound T_k = |2 pi (d_k - D_k)| max|r'|; beyond T / sqrt 2 its sums
        // take the factor 1 / (1 + T_k^2 / 2 - T^2 / 4) (the planner's truncation bound)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const float t2 = (d[0][k][0] * d[0][k][0] + d[0][k][1] * d[0][k][1] + d[0][k][2] * d[0][k][2]) *
                             (g.refrho * g.refrho);
            const float cs = fmaxf(0.f, 0.5f * t2 - g.refc);
            if (cs > 0.f) {
                const float sc = 1.f / (1.f + cs);
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    acc[0][k][m] *= sc;
                    cmp[0][k][m] *= sc;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// FP32-split tile with four mirror pairs about a 2 x 4 block (tile variant 8; the
// 16 x 16 lanes cover 32 x 64 pixels).  Pair k joins pixel (0, k) (first) and
// (1, 3 - k) (second) of the lane's block, both at offset o_k from the block
// centre.  Every pair has its own centre m_k (its mean dq), so nothing is
// dropped; one sincos at pair 1's centre (|o_1| smallest) serves all four, the
// others reached by the first-order rotation e^{-i Delta_k} ~ (1 - i Delta_k),
// Delta_k = 2 pi (m_1 - m_k) . r'.  All pairs share the block centre, so the
// m_k differ only by second-order terms and Delta_k stays a few 1e-3 rad at
// most (the host planner bounds Delta_k^2 / 2).  Per 8 pixels: 10
// transcendentals and 46 FMA.  The W term must be folded into A.
template <bool HAS_AMP>
__device__ __forceinline__ void block_tile(const Geom& g, const double* __restrict__ emit,
                                           const double* __restrict__ amp, int64_t e0, int64_t e1, int tx0, int ty0,
                                           float4* s0, float2* s1, float (&acc)[4][4], float (&cmp)[4][4])
{
    constexpr int TX = 32, TY = 64;
    const int tid = threadIdx.x;
    const int lx = tid >> 4, ly = tid & 15;

    double ccx, ccy;
    tile_centre(g, tx0, ty0, TX, TY, ccx, ccy);
    const Frame c = pixel_frame(g, ccx, ccy);

    float mc[3], d[4][3], dv[4][3];
    {
        double dq[2][4][3];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const Frame f = pixel_frame(g, (double)(tx0 + 2 * lx + i), (double)(ty0 + 4 * ly + j));
#pragma unroll
                for (int a = 0; a < 3; ++a) dq[i][j][a] = -g.krev * (f.u[a] - c.u[a]);
            }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            double m[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                m[k] = 0.5 * (dq[0][k][a] + dq[1][3 - k][a]);
                d[k][a] = (float)(0.5 * (dq[1][3 - k][a] - dq[0][k][a]));
            }
            mc[a] = (float)m[1];
#pragma unroll
            for (int k = 0; k < 4; ++k) dv[k][a] = (float)(kTwoPi * (m[1] - m[k]));
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            acc[k][n] = 0.f;
            cmp[k][n] = 0.f;
        }
    }

    for (int64_t base = e0; base < e1; base += kChunk) {
        const int cnt = (e1 - base) < (int64_t)kChunk ? (int)(e1 - base) : kChunk;
        __syncthreads();  // the previous chunk has been consumed by every lane
        if (tid < cnt) {
            const int64_t e = base + tid;
            const double* a4 = emit + 4 * e;
            const double ea = (HAS_AMP && amp != nullptr) ? amp[e] : 1.0;
            const SplitRec r = make_split_rec(g, c, a4[0], a4[1], a4[2], a4[3], ea);
            s0[tid] = make_float4(r.x, r.y, r.z, r.A);
            s1[tid] = make_float2(r.W, r.amp);
        }
        __syncthreads();

        float su[4][4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int n = 0; n < 4; ++n) su[k][n] = 0.f;
#pragma unroll 2
        for (int e = 0; e < cnt; ++e) {
            const float4 r0 = s0[e];
            float t = fmaf(mc[0], r0.x, r0.w);
            t = fmaf(mc[1], r0.y, t);
            t = fmaf(mc[2], r0.z, t);
            float sc = __builtin_amdgcn_sinf(t);
            float cc = __builtin_amdgcn_cosf(t);
            if (HAS_AMP) {
                const float am = s1[e].y;
                sc *= am;
                cc *= am;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                float pc = cc, ps = sc;
                if (k != 1) {
                    float dl = dv[k][0] * r0.x;
                    dl = fmaf(dv[k][1], r0.y, dl);
                    dl = fmaf(dv[k][2], r0.z, dl);
                    pc = fmaf(dl, sc, cc);  // (cc + i sc)(1 - i dl)
                    ps = fmaf(-dl, cc, sc);
                }
                float dl = d[k][0] * r0.x;
                dl = fmaf(d[k][1], r0.y, dl);
                dl = fmaf(d[k][2], r0.z, dl);
                const float sd = __builtin_amdgcn_sinf(dl);
                const float cd = __builtin_amdgcn_cosf(dl);
                su[k][0] = fmaf(pc, cd, su[k][0]);
                su[k][1] = fmaf(ps, cd, su[k][1]);
                su[k][2] = fmaf(pc, sd, su[k][2]);
                su[k][3] = fmaf(ps, sd, su[k][3]);
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int n = 0; n < 4; ++n) kahan_add(acc[k][n], cmp[k][n], su[k][n]);
    }
}

// Pixel (i, j) of a lane's 2 x 4 block: pair k = j (first member, i = 0) or 3 - j (second, i = 1).
__device__ __forceinline__ void block_pixel(const float (&acc)[4][4], const float (&cmp)[4][4], int i, int j,
                                            double& re, double& im)
{
    const int k = (i == 0) ? j : 3 - j;
    const double ur = (double)acc[k][0] - (double)cmp[k][0];
    const double ui = (double)acc[k][1] - (double)cmp[k][1];
    const double vr = (double)acc[k][2] - (double)cmp[k][2];
    const double vi = (double)acc[k][3] - (double)cmp[k][3];
    re = (i == 1) ? ur - vi : ur + vi;
    im = (i == 1) ? ui + vr : ui - vr;
}

// The four pixels of quad q of a lane: pixel (i, j) of the quad, its field in FP64
// from the compensated U and V sums of its pair.
template <int QY>
__device__ __forceinline__ void pair_pixel(const float (&acc)[QY][2][4], const float (&cmp)[QY][2][4], int q,
                                           int i, int j, double& re, double& im)
{
    // pair 0: (0,0) first, (1,1) second; pair 1: (1,0) first, (0,1) second
    const int k = (i == j) ? 0 : 1;
    const bool second = (k == 0) ? (i == 1) : (i == 0);
    const double ur = (double)acc[q][k][0] - (double)cmp[q][k][0];
    const double ui = (double)acc[q][k][1] - (double)cmp[q][k][1];
    const double vr = (double)acc[q][k][2] - (double)cmp[q][k][2];
    const double vi = (double)acc[q][k][3] - (double)cmp[q][k][3];
    // (U +/- i V): re = ur -/+ vi, im = ui +/- vr
    re = second ? ur - vi : ur + vi;
    im = second ? ui + vr : ui - vr;
}

// An FP64 value as the FP32 pair (hi, lo) with hi - lo ~ v, the form of the
// split partials (the finalize kernel forms hi - lo in FP64).
__device__ __forceinline__ void split_hi_lo(double v, float& hi, float& lo)
{
    hi = (float)v;
    lo = (float)((double)hi - v);
}

// tab: the tiles' pair_setup tables (k_pair_table; QY 1), or nullptr to form them in the prologue.
template <int QY, bool HAS_AMP, bool CENTRED, int REF>
__global__ __launch_bounds__(kBlock) void k_fwd_pair(Geom g, const double* __restrict__ emit,
                                                     const double* __restrict__ amp, int64_t n, int tiles_y,
                                                     int64_t per_split, const float* __restrict__ tab,
                                                     float4* __restrict__ partial, void* out, int kind)
{
    constexpr int TX = 32, TY = 32 * QY;
    __shared__ float4 s0[pair_chunk(REF)];
    __shared__ float2 s1[pair_s1_float2(REF)];
    __shared__ float sD[pair_sd_floats(REF)];
    __shared__ float4 sE[pair_se_float4(REF)];
    __shared__ float sC[pair_sc_floats(REF)];
    const int tile = blockIdx.x;
    const int split = blockIdx.y;
    const int tx0 = (tile / tiles_y) * TX, ty0 = (tile % tiles_y) * TY;
    const int64_t e0 = (int64_t)split * per_split;
    const int64_t e1 = (e0 + per_split < n) ? e0 + per_split : n;
    float acc[QY][2][4], cmp[QY][2][4];
    const float* tt = tab != nullptr ? tab + (size_t)tile * pair_tab_floats(REF) : nullptr;
    pair_tile<QY, HAS_AMP, CENTRED, REF, true>(g, emit, amp, e0, e1, tx0, ty0, tt, s0, s1, sD, sE, sC, acc, cmp);

    int lx, ly;
    pair_lane<REF>(threadIdx.x, lx, ly);
    const Geom gs = item_geom();  // store frames formed here (see k_batched_pair)
    const size_t npix = (size_t)gs.n0 * gs.n1;
#pragma unroll
    for (int q = 0; q < QY; ++q) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int x = tx0 + 2 * lx + i, y = ty0 + 2 * QY * ly + 2 * q + j;
                if (x < gs.n0 && y < gs.n1) {
                    double re, im;
                    pair_pixel<QY>(acc, cmp, q, i, j, re, im);
                    const size_t idx = (size_t)x * gs.n1 + y;
                    if (partial != nullptr) {
                        float rh, rl, ih, il;
                        split_hi_lo(re, rh, rl);
                        split_hi_lo(im, ih, il);
                        partial[(size_t)split * npix + idx] = make_float4(rh, ih, rl, il);
                    } else {
                        store_split(gs, x, y, re, im, out, kind, idx);
                    }
                }
            }
        }
    }
}

// The pair_setup table of every 32 x 32 pixel tile (one quad per lane): one workgroup per tile.
template <bool CENTRED, int REF>
__global__ __launch_bounds__(kBlock) void k_pair_table(Geom g, int tiles_y, float* __restrict__ tab)
{
    __shared__ float sD[pair_sd_floats(REF)];
    __shared__ float4 sE[REF == 3 ? 60 : 1];  // REF 3: the half differences, 4 waves x 5 lanes x 6 doubles
    const int tile = blockIdx.x, tid = threadIdx.x;
    const int tx0 = (tile / tiles_y) * 32, ty0 = (tile % tiles_y) * 32;
    for (int i = tid; i < pair_sd_floats(REF); i += kBlock) sD[i] = 0.f;
    __syncthreads();
    int lx, ly;
    pair_lane<REF>(tid, lx, ly);
    const int myref = (REF == 4 || REF == 5) ? pair_subblock(tid) : 0;
    Frame c;
    float mc[1][3], dv[1][3], d[1][2][3];
    pair_setup<1, CENTRED, REF>(g, tx0, ty0, tid, lx, ly, myref, sD, sE, c, mc, dv, d);
    float* t = tab + (size_t)tile * pair_tab_floats(REF);
    float4* tl = reinterpret_cast<float4*>(t);
    tl[tid] = make_float4(mc[0][0], mc[0][1], mc[0][2], dv[0][0]);
    tl[kBlock + tid] = make_float4(dv[0][1], dv[0][2], d[0][0][0], d[0][0][1]);
    tl[2 * kBlock + tid] = make_float4(d[0][0][2], d[0][1][0], d[0][1][1], d[0][1][2]);
    __syncthreads();  // sD complete
    for (int i = tid; i < pair_tab_sd(REF); i += kBlock) t[kPairTabLane + i] = i < pair_sd_floats(REF) ? sD[i] : 0.f;
    if (tid == 0) {
        double* fr = reinterpret_cast<double*>(t + kPairTabLane + pair_tab_sd(REF));
        fr[0] = c.R;
        fr[1] = c.rh[0], fr[2] = c.rh[1], fr[3] = c.rh[2];
        fr[4] = c.u[0], fr[5] = c.u[1], fr[6] = c.u[2];
        fr[7] = c.RmZ;
    }
}

template <bool HAS_AMP>
__global__ __launch_bounds__(kBlock) void k_fwd_block(Geom g, const double* __restrict__ emit,
                                                      const double* __restrict__ amp, int64_t n, int tiles_y,
                                                      int64_t per_split, float4* __restrict__ partial, void* out,
                                                      int kind)
{
    constexpr int TX = 32, TY = 64;
    __shared__ float4 s0[kChunk];
    __shared__ float2 s1[kChunk];
    const int tile = blockIdx.x;
    const int split = blockIdx.y;
    const int tx0 = (tile / tiles_y) * TX, ty0 = (tile % tiles_y) * TY;
    const int64_t e0 = (int64_t)split * per_split;
    const int64_t e1 = (e0 + per_split < n) ? e0 + per_split : n;
    float acc[4][4], cmp[4][4];
    block_tile<HAS_AMP>(g, emit, amp, e0, e1, tx0, ty0, s0, s1, acc, cmp);

    const int lx = threadIdx.x >> 4, ly = threadIdx.x & 15;
    const size_t npix = (size_t)g.n0 * g.n1;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int x = tx0 + 2 * lx + i, y = ty0 + 4 * ly + j;
            if (x < g.n0 && y < g.n1) {
                double re, im;
                block_pixel(acc, cmp, i, j, re, im);
                const size_t idx = (size_t)x * g.n1 + y;
                if (partial != nullptr) {
                    float rh, rl, ih, il;
                    split_hi_lo(re, rh, rl);
                    split_hi_lo(im, ih, il);
                    partial[(size_t)split * npix + idx] = make_float4(rh, ih, rl, il);
                } else {
                    store_split(g, x, y, re, im, out, kind, idx);
                }
            }
        }
    }
}

template <bool HAS_AMP>
__global__ __launch_bounds__(kBlock) void k_batched_block(Geom g0, const double* __restrict__ emit,
                                                          const double* __restrict__ amp,
                                                          const WorkItem* __restrict__ items, int nitems,
                                                          int* __restrict__ counter, int tiles_y, void* out, int kind)
{
    constexpr int TX = 32, TY = 64;
    __shared__ float4 s0[kChunk];
    __shared__ float2 s1[kChunk];
    __shared__ int s_item;
    const int lx = threadIdx.x >> 4, ly = threadIdx.x & 15;
    const size_t npix = (size_t)g0.n0 * g0.n1;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = s_item;
        if (it >= nitems) break;
        const WorkItem w = items[it];
        const Geom g = item_geom();
        const int tx0 = (w.tile / tiles_y) * TX, ty0 = (w.tile % tiles_y) * TY;
        float acc[4][4], cmp[4][4];
        block_tile<HAS_AMP>(g, emit, amp, w.e0, w.e1, tx0, ty0, s0, s1, acc, cmp);
        char* outs = static_cast<char*>(out) +
                     (size_t)w.shot * npix * (kind == OUT_INTENSITY ? sizeof(float) : sizeof(float2));
        const Geom gs = item_geom();  // store frames formed here (see k_batched_pair)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int x = tx0 + 2 * lx + i, y = ty0 + 4 * ly + j;
                if (x < gs.n0 && y < gs.n1) {
                    double re, im;
                    block_pixel(acc, cmp, i, j, re, im);
                    store_split(gs, x, y, re, im, outs, kind, (size_t)x * gs.n1 + y);
                }
            }
        }
    }
}

// Persistent batched form of k_fwd_pair (the queue protocol of k_batched_split).  The one-quad
// variants are held to 128 VGPRs (four waves per SIMD, as their LDS allows): without the bound
// the allocator settles a few registers above it and one wave per SIMD is lost.
template <int QY, bool HAS_AMP, bool CENTRED, int REF>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(QY == 1 ? 4 : 2))) void k_batched_pair(Geom g0, const double* __restrict__ emit,
                                                         const double* __restrict__ amp,
                                                         const WorkItem* __restrict__ items, int nitems,
                                                         int* __restrict__ counter, int tiles_y, void* out, int kind)
{
    constexpr int TX = 32, TY = 32 * QY;
    __shared__ float4 s0[pair_chunk(REF)];
    __shared__ float2 s1[pair_s1_float2(REF)];
    __shared__ float sD[pair_sd_floats(REF)];
    __shared__ float4 sE[pair_se_float4(REF)];
    __shared__ float sC[pair_sc_floats(REF)];
    __shared__ int s_item;
    int lx, ly;
    pair_lane<REF>(threadIdx.x, lx, ly);
    const size_t npix = (size_t)g0.n0 * g0.n1;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = s_item;
        if (it >= nitems) break;
        const WorkItem w = items[it];
        const Geom g = item_geom();
        const int tx0 = (w.tile / tiles_y) * TX, ty0 = (w.tile % tiles_y) * TY;
        float acc[QY][2][4], cmp[QY][2][4];
        pair_tile<QY, HAS_AMP, CENTRED, REF>(g, emit, amp, w.e0, w.e1, tx0, ty0, nullptr, s0, s1, sD, sE, sC, acc,
                                             cmp);
        char* outs = static_cast<char*>(out) +
                     (size_t)w.shot * npix * (kind == OUT_INTENSITY ? sizeof(float) : sizeof(float2));
        // read again, so that the pixel frames of the stores are formed here and not kept in
        // VGPRs from the prologue's identical frames across the emitter loop
        const Geom gs = item_geom();
#pragma unroll
        for (int q = 0; q < QY; ++q) {
#pragma unroll
            for (int i = 0; i < 2; ++i) {
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const int x = tx0 + 2 * lx + i, y = ty0 + 2 * QY * ly + 2 * q + j;
                    if (x < gs.n0 && y < gs.n1) {
                        double re, im;
                        pair_pixel<QY>(acc, cmp, q, i, j, re, im);
                        store_split(gs, x, y, re, im, outs, kind, (size_t)x * gs.n1 + y);
                    }
                }
            }
        }
    }
}

// FP64 combine of the split partials: a workgroup owns 64 pixels, its four rows of
// threads stride over the splits (several loads in flight per pixel) and the rows are
// summed through LDS.
__global__ __launch_bounds__(256) void k_finalize_split(Geom g, const float4* __restrict__ partial, int splits,
                                                        void* out, int kind)
{
    __shared__ double red[2][4][64];
    const size_t npix = (size_t)g.n0 * g.n1;
    const int col = threadIdx.x & 63, row = threadIdx.x >> 6;
    const size_t idx = (size_t)blockIdx.x * 64 + col;
    double re0 = 0.0, im0 = 0.0, re1 = 0.0, im1 = 0.0;
    if (idx < npix) {
        int s = row;
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
    }
    red[0][row][col] = re0 + re1;
    red[1][row][col] = im0 + im1;
    __syncthreads();
    if (row == 0 && idx < npix) {
        const double re = (red[0][0][col] + red[0][1][col]) + (red[0][2][col] + red[0][3][col]);
        const double im = (red[1][0][col] + red[1][1][col]) + (red[1][2][col] + red[1][3][col]);
        store_split(g, (int)(idx / g.n1), (int)(idx % g.n1), re, im, out, kind, idx);
    }
}

// ---------------------------------------------------------------------------
// Literal FP64 kernel (cpu.py:42-105 / sim.cu:43-139, FP64 geometry).
template <int PX, int PY>
__device__ __forceinline__ void fp64_tile(const Geom& g, const double* __restrict__ emit,
                                          const double* __restrict__ amp, int64_t e0, int64_t e1,
                                          int tx0, int ty0, double4* s0, double* s1,
                                          double (&acc_re)[PX * PY], double (&acc_im)[PX * PY],
                                          double (&cmp_re)[PX * PY], double (&cmp_im)[PX * PY])
{
    constexpr int P = PX * PY;
    const int tid = threadIdx.x;
    const int lx = tid >> 4, ly = tid & 15;
    double dx[P], dy[P], dz[P], qx[P], qy[P], qz[P], rd[P], c2[P];
#pragma unroll
    for (int i = 0; i < PX; ++i) {
#pragma unroll
        for (int j = 0; j < PY; ++j) {
            const int p = i * PY + j;
            double r[3];
            pixel_pos(g, (double)(tx0 + lx * PX + i), (double)(ty0 + ly * PY + j), r);
            dx[p] = r[0];
            dy[p] = r[1];
            dz[p] = r[2];
            // cpu.py:85-89 / sim.cu:77-84
            rd[p] = 1.0 / sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            qz[p] = g.k * (r[2] * rd[p] - 1.0);
            qx[p] = g.k * (r[0] * rd[p]);
            qy[p] = g.k * (r[1] * rd[p]);
            c2[p] = g.k * rd[p] / 2.0;
            acc_re[p] = 0.0;
            acc_im[p] = 0.0;
            cmp_re[p] = 0.0;
            cmp_im[p] = 0.0;
        }
    }
    for (int64_t base = e0; base < e1; base += kChunk) {
        const int cnt = (e1 - base) < (int64_t)kChunk ? (int)(e1 - base) : kChunk;
        __syncthreads();
        if (tid < cnt) {
            const int64_t e = base + tid;
            const double* a4 = emit + 4 * e;
            s0[tid] = make_double4(a4[0], a4[1], a4[2], a4[3]);
            s1[tid] = (amp != nullptr) ? amp[e] : 1.0;
        }
        __syncthreads();
        // one emitter loop per mode (near or far field, with or without the 1/r amplitude): the
        // mode is uniform, and a loop that selects inside was if-converted to compute both phases
        // (and the square root) for every pair
        auto emitters = [&](auto nf_tag, auto scale_tag) {
            constexpr bool NF = decltype(nf_tag)::value, SCALE = decltype(scale_tag)::value;
            for (int e = 0; e < cnt; ++e) {
                const double4 at = s0[e];
                const double a = s1[e];
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    double phase, mag;
                    if constexpr (NF) {
                        const double ddx = dx[p] - at.x, ddy = dy[p] - at.y, ddz = dz[p] - at.z;
                        const double dist = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
                        phase = (dist - dz[p] + at.z) * g.k + at.w;
                        mag = SCALE ? a / dist : a;
                    } else {
                        phase = -(at.x * qx[p] + at.y * qy[p] + at.z * qz[p]) + at.w;
                        if (g.secondorder) phase += c2[p] * (at.x * at.x + at.y * at.y);
                        mag = SCALE ? a * rd[p] : a;
                    }
                    double sn, cs;
                    sincos_fp64(phase, &sn, &cs);
                    kahan_add_d(acc_re[p], cmp_re[p], mag * cs);
                    kahan_add_d(acc_im[p], cmp_im[p], mag * sn);
                }
            }
        };
        if (g.nf) {
            if (g.scale) emitters(std::true_type{}, std::true_type{});
            else emitters(std::true_type{}, std::false_type{});
        } else {
            if (g.scale) emitters(std::false_type{}, std::true_type{});
            else emitters(std::false_type{}, std::false_type{});
        }
    }
}

template <int PX, int PY>
__global__ __launch_bounds__(kBlock) void k_fwd_fp64(Geom g, const double* __restrict__ emit,
                                                     const double* __restrict__ amp, int64_t n, int tiles_y,
                                                     int64_t per_split, double4* __restrict__ partial,
                                                     void* out, int kind)
{
    constexpr int P = PX * PY;
    constexpr int TX = 16 * PX, TY = 16 * PY;
    __shared__ double4 s0[kChunk];
    __shared__ double s1[kChunk];
    const int tile = blockIdx.x, split = blockIdx.y;
    const int tx0 = (tile / tiles_y) * TX, ty0 = (tile % tiles_y) * TY;
    const int64_t e0 = (int64_t)split * per_split;
    const int64_t e1 = (e0 + per_split < n) ? e0 + per_split : n;
    double re[P], im[P], cre[P], cim[P];
    fp64_tile<PX, PY>(g, emit, amp, e0, e1, tx0, ty0, s0, s1, re, im, cre, cim);
    const int lx = threadIdx.x >> 4, ly = threadIdx.x & 15;
    const size_t npix = (size_t)g.n0 * g.n1;
#pragma unroll
    for (int i = 0; i < PX; ++i) {
#pragma unroll
        for (int j = 0; j < PY; ++j) {
            const int p = i * PY + j;
            const int x = tx0 + lx * PX + i, y = ty0 + ly * PY + j;
            if (x < g.n0 && y < g.n1) {
                const size_t idx = (size_t)x * g.n1 + y;
                if (partial != nullptr)
                    partial[(size_t)split * npix + idx] = make_double4(re[p], im[p], cre[p], cim[p]);
                else
                    store_fp64(kind, re[p] - cre[p], im[p] - cim[p], out, idx);
            }
        }
    }
}

template <int PX, int PY>
__global__ __launch_bounds__(kBlock) void k_batched_fp64(Geom g0, const double* __restrict__ emit,
                                                         const double* __restrict__ amp,
                                                         const WorkItem* __restrict__ items, int nitems,
                                                         int* __restrict__ counter, int tiles_y, void* out,
                                                         int kind)
{
    constexpr int P = PX * PY;
    constexpr int TX = 16 * PX, TY = 16 * PY;
    __shared__ double4 s0[kChunk];
    __shared__ double s1[kChunk];
    __shared__ int s_item;
    const int lx = threadIdx.x >> 4, ly = threadIdx.x & 15;
    const size_t npix = (size_t)g0.n0 * g0.n1;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = s_item;
        if (it >= nitems) break;
        const WorkItem w = items[it];
        const Geom g = item_geom();
        const int tx0 = (w.tile / tiles_y) * TX, ty0 = (w.tile % tiles_y) * TY;
        double re[P], im[P], cre[P], cim[P];
        fp64_tile<PX, PY>(g, emit, amp, w.e0, w.e1, tx0, ty0, s0, s1, re, im, cre, cim);
        char* outs = static_cast<char*>(out) +
                     (size_t)w.shot * npix * (kind == OUT_INTENSITY ? sizeof(double) : sizeof(double2));
#pragma unroll
        for (int i = 0; i < PX; ++i) {
#pragma unroll
            for (int j = 0; j < PY; ++j) {
                const int p = i * PY + j;
                const int x = tx0 + lx * PX + i, y = ty0 + ly * PY + j;
                if (x < g.n0 && y < g.n1)
                    store_fp64(kind, re[p] - cre[p], im[p] - cim[p], outs, (size_t)x * g.n1 + y);
            }
        }
    }
}

// kind bit 1 set: write complex64 / float32 (FP64 fallback of an FP32-split handle)
__global__ void k_finalize_fp64(Geom g, const double4* __restrict__ partial, int splits, void* out,
                                int kind)
{
    const size_t npix = (size_t)g.n0 * g.n1;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= npix) return;
    double re = 0.0, im = 0.0, cre = 0.0, cim = 0.0;
    for (int s = 0; s < splits; ++s) {
        const double4 v = partial[(size_t)s * npix + idx];
        kahan_add_d(re, cre, v.x - v.z);
        kahan_add_d(im, cim, v.y - v.w);
    }
    re -= cre;
    im -= cim;
    if (kind & 2) {
        if ((kind & 1) == OUT_INTENSITY)
            static_cast<float*>(out)[idx] = (float)(re * re + im * im);
        else
            static_cast<float2*>(out)[idx] = make_float2((float)re, (float)im);
    } else {
        store_fp64(kind, re, im, out, idx);
    }
}

// ---------------------------------------------------------------------------
__global__ void k_reset_counter(int* counter) { *counter = 0; }

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v) { *p = v; }

// Zero `words` 32-bit words (a kernel, so that a captured graph holds no memset node).
__global__ __launch_bounds__(256) void k_zero_words(uint32_t* __restrict__ p, int64_t words)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// max_j |r_j|^2 as the bit pattern of a non-negative double (ordered like u64)
__global__ __launch_bounds__(kBlock) void k_extent(const double* __restrict__ emit, int64_t n,
                                                   unsigned long long* out)
{
    __shared__ double red[kBlock];
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double* a = emit + 4 * i;
        m = fmax(m, a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(red[0]));
}

// ---------------------------------------------------------------------------
// Launchers.
template <int V, bool W, bool A, int S>
static hipError_t launch_split_v(const Geom& g, const double* emit, const double* amp, int64_t n, int splits,
                                 float4* partial, void* out, int kind, hipStream_t st)
{
    constexpr int PX = tile_px(V), PY = tile_py(V);
    const int tiles_x = (g.n0 + 16 * PX - 1) / (16 * PX);
    const int tiles_y = (g.n1 + 16 * PY - 1) / (16 * PY);
    const int64_t per = split_size(n, splits);
    dim3 grid(tiles_x * tiles_y, splits);
    hipLaunchKernelGGL((k_fwd_split<PX, PY, W, A, S>), grid, dim3(kBlock), 0, st, g, emit, amp, n, tiles_y,
                       per, partial, out, kind);
    return hipGetLastError();
}

template <int V, bool W, bool A, int S>
static hipError_t launch_batched_split_v(const Geom& g, const double* emit, const double* amp,
                                         const WorkItem* items, int nitems, int* counter, int nblocks,
                                         void* out, int kind, hipStream_t st)
{
    constexpr int PX = tile_px(V), PY = tile_py(V);
    const int tiles_y = (g.n1 + 16 * PY - 1) / (16 * PY);
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, st, counter);
    hipLaunchKernelGGL((k_batched_split<PX, PY, W, A, S>), dim3(nblocks), dim3(kBlock), 0, st, g, emit, amp,
                       items, nitems, counter, tiles_y, out, kind);
    return hipGetLastError();
}

template <int V, bool W, bool A, int S>
static int occupancy_split_v()
{
    constexpr int PX = tile_px(V), PY = tile_py(V);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batched_split<PX, PY, W, A, S>, kBlock, 0) !=
        hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

#define IDISIM_DISPATCH_SPLIT(FN, ...)                                                        \
    switch ((v << 3) | ((int)w << 2) | ((int)a << 1) | s) {                                   \
        case 0x00: return FN<0, false, false, 0>(__VA_ARGS__);                                \
        case 0x01: return FN<0, false, false, 1>(__VA_ARGS__);                                \
        case 0x02: return FN<0, false, true, 0>(__VA_ARGS__);                                 \
        case 0x03: return FN<0, false, true, 1>(__VA_ARGS__);                                 \
        case 0x04: return FN<0, true, false, 0>(__VA_ARGS__);                                 \
        case 0x05: return FN<0, true, false, 1>(__VA_ARGS__);                                 \
        case 0x06: return FN<0, true, true, 0>(__VA_ARGS__);                                  \
        case 0x07: return FN<0, true, true, 1>(__VA_ARGS__);                                  \
        case 0x08: return FN<1, false, false, 0>(__VA_ARGS__);                                \
        case 0x09: return FN<1, false, false, 1>(__VA_ARGS__);                                \
        case 0x0a: return FN<1, false, true, 0>(__VA_ARGS__);                                 \
        case 0x0b: return FN<1, false, true, 1>(__VA_ARGS__);                                 \
        case 0x0c: return FN<1, true, false, 0>(__VA_ARGS__);                                 \
        case 0x0d: return FN<1, true, false, 1>(__VA_ARGS__);                                 \
        case 0x0e: return FN<1, true, true, 0>(__VA_ARGS__);                                  \
        case 0x0f: return FN<1, true, true, 1>(__VA_ARGS__);                                  \
        case 0x10: return FN<2, false, false, 0>(__VA_ARGS__);                                \
        case 0x11: return FN<2, false, false, 1>(__VA_ARGS__);                                \
        case 0x12: return FN<2, false, true, 0>(__VA_ARGS__);                                 \
        case 0x13: return FN<2, false, true, 1>(__VA_ARGS__);                                 \
        case 0x14: return FN<2, true, false, 0>(__VA_ARGS__);                                 \
        case 0x15: return FN<2, true, false, 1>(__VA_ARGS__);                                 \
        case 0x16: return FN<2, true, true, 0>(__VA_ARGS__);                                  \
        case 0x17: return FN<2, true, true, 1>(__VA_ARGS__);                                  \
        case 0x18: return FN<3, false, false, 0>(__VA_ARGS__);                                \
        case 0x19: return FN<3, false, false, 1>(__VA_ARGS__);                                \
        case 0x1a: return FN<3, false, true, 0>(__VA_ARGS__);                                 \
        case 0x1b: return FN<3, false, true, 1>(__VA_ARGS__);                                 \
        case 0x1c: return FN<3, true, false, 0>(__VA_ARGS__);                                 \
        case 0x1d: return FN<3, true, false, 1>(__VA_ARGS__);                                 \
        case 0x1e: return FN<3, true, true, 0>(__VA_ARGS__);                                  \
        case 0x1f: return FN<3, true, true, 1>(__VA_ARGS__);                                  \
        default: break;                                                                       \
    }

hipError_t launch_fwd_split(int v, bool w, bool a, int s, const Geom& g, const double* emit, const double* amp,
                            int64_t n, int splits, float4* partial, void* out, int kind, hipStream_t st)
{
    IDISIM_DISPATCH_SPLIT(launch_split_v, g, emit, amp, n, splits, partial, out, kind, st)
    return hipErrorInvalidValue;
}

hipError_t launch_batched_split(int v, bool w, bool a, int s, const Geom& g, const double* emit,
                                const double* amp, const WorkItem* items, int nitems, int* counter, int nblocks,
                                void* out, int kind, hipStream_t st)
{
    IDISIM_DISPATCH_SPLIT(launch_batched_split_v, g, emit, amp, items, nitems, counter, nblocks, out, kind, st)
    return hipErrorInvalidValue;
}

int occupancy_batched_split(int v, bool w, bool a, int s)
{
    IDISIM_DISPATCH_SPLIT(occupancy_split_v)
    return 1;
}

template <int QY, bool A, bool C, int R>
static hipError_t launch_pair_q(const Geom& g, const double* emit, const double* amp, int64_t n, int splits,
                                const float* tab, float4* partial, void* out, int kind, hipStream_t st)
{
    const int tiles_x = (g.n0 + 31) / 32;
    const int tiles_y = (g.n1 + 32 * QY - 1) / (32 * QY);
    const int64_t per = split_size(n, splits);
    dim3 grid(tiles_x * tiles_y, splits);
    hipLaunchKernelGGL((k_fwd_pair<QY, A, C, R>), grid, dim3(kBlock), 0, st, g, emit, amp, n, tiles_y,
                       per, QY == 1 ? tab : nullptr, partial, out, kind);
    return hipGetLastError();
}

template <int QY, bool A, bool C, int R>
static hipError_t launch_pair_table_q(const Geom& g, float* tab, hipStream_t st)
{
    if (QY != 1) return hipErrorInvalidValue;
    const int tiles_x = (g.n0 + 31) / 32, tiles_y = (g.n1 + 31) / 32;
    hipLaunchKernelGGL((k_pair_table<C, R>), dim3(tiles_x * tiles_y), dim3(kBlock), 0, st, g, tiles_y, tab);
    return hipGetLastError();
}

template <int QY, bool A, bool C, int R>
static hipError_t launch_batched_pair_q(const Geom& g, const double* emit, const double* amp,
                                        const WorkItem* items, int nitems, int* counter, int nblocks, void* out,
                                        int kind, hipStream_t st)
{
    const int tiles_y = (g.n1 + 32 * QY - 1) / (32 * QY);
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, st, counter);
    hipLaunchKernelGGL((k_batched_pair<QY, A, C, R>), dim3(nblocks), dim3(kBlock), 0, st, g, emit, amp, items, nitems,
                       counter, tiles_y, out, kind);
    return hipGetLastError();
}

template <int QY, bool A, bool C, int R>
static int occupancy_pair_q()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batched_pair<QY, A, C, R>, kBlock, 0) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

// (quads per lane, reference kind, centred, amplitudes) -> instantiation; only
// variant 6 has two quads per lane
#define IDISIM_DISPATCH_PAIR(FN, ...)                                                          \
    switch ((pair_quads(v) << 5) | (tile_ref_kind(v) << 2) | ((int)tile_pair_centred(v) << 1) | (int)a) { \
        case 32: return FN<1, false, false, 0>(__VA_ARGS__);                                  \
        case 33: return FN<1, true, false, 0>(__VA_ARGS__);                                   \
        case 34: return FN<1, false, true, 0>(__VA_ARGS__);                                   \
        case 35: return FN<1, true, true, 0>(__VA_ARGS__);                                    \
        case 36: return FN<1, false, false, 1>(__VA_ARGS__);                                  \
        case 37: return FN<1, true, false, 1>(__VA_ARGS__);                                   \
        case 38: return FN<1, false, true, 1>(__VA_ARGS__);                                   \
        case 39: return FN<1, true, true, 1>(__VA_ARGS__);                                    \
        case 40: return FN<1, false, false, 2>(__VA_ARGS__);                                  \
        case 41: return FN<1, true, false, 2>(__VA_ARGS__);                                   \
        case 42: return FN<1, false, true, 2>(__VA_ARGS__);                                   \
        case 43: return FN<1, true, true, 2>(__VA_ARGS__);                                    \
        case 44: return FN<1, false, false, 3>(__VA_ARGS__);                                  \
        case 45: return FN<1, true, false, 3>(__VA_ARGS__);                                   \
        case 50: return FN<1, false, true, 4>(__VA_ARGS__);                                   \
        case 51: return FN<1, true, true, 4>(__VA_ARGS__);                                    \
        case 54: return FN<1, false, true, 5>(__VA_ARGS__);                                   \
        case 55: return FN<1, true, true, 5>(__VA_ARGS__);                                    \
        case 64: return FN<2, false, false, 0>(__VA_ARGS__);                                  \
        case 65: return FN<2, true, false, 0>(__VA_ARGS__);                                   \
        default: break;                                                                       \
    }

template <bool A>
static hipError_t launch_block_a(const Geom& g, const double* emit, const double* amp, int64_t n, int splits,
                                 float4* partial, void* out, int kind, hipStream_t st)
{
    const int tiles_x = (g.n0 + 31) / 32;
    const int tiles_y = (g.n1 + 63) / 64;
    const int64_t per = split_size(n, splits);
    dim3 grid(tiles_x * tiles_y, splits);
    hipLaunchKernelGGL((k_fwd_block<A>), grid, dim3(kBlock), 0, st, g, emit, amp, n, tiles_y,
                       per, partial, out, kind);
    return hipGetLastError();
}

template <bool A>
static hipError_t launch_batched_block_a(const Geom& g, const double* emit, const double* amp,
                                         const WorkItem* items, int nitems, int* counter, int nblocks, void* out,
                                         int kind, hipStream_t st)
{
    const int tiles_y = (g.n1 + 63) / 64;
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, st, counter);
    hipLaunchKernelGGL((k_batched_block<A>), dim3(nblocks), dim3(kBlock), 0, st, g, emit, amp, items, nitems,
                       counter, tiles_y, out, kind);
    return hipGetLastError();
}

template <bool A>
static int occupancy_block_a()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batched_block<A>, kBlock, 0) != hipSuccess) return 1;
    return nb > 0 ? nb : 1;
}

hipError_t launch_fwd_pair(int v, bool a, const Geom& g, const double* emit, const double* amp, int64_t n,
                           int splits, const float* tab, float4* partial, void* out, int kind, hipStream_t st)
{
    if (g.foldw == 0 && (g.nf || g.secondorder)) return hipErrorInvalidValue;  // needs W folded into A
    if (v == kTileBlock)
        return a ? launch_block_a<true>(g, emit, amp, n, splits, partial, out, kind, st)
                 : launch_block_a<false>(g, emit, amp, n, splits, partial, out, kind, st);
    IDISIM_DISPATCH_PAIR(launch_pair_q, g, emit, amp, n, splits, tab, partial, out, kind, st)
    return hipErrorInvalidValue;
}

size_t pair_table_bytes(int v, const Geom& g)
{
    if (!tile_pair(v) || v == kTileBlock || pair_quads(v) != 1) return 0;
    const size_t tiles = (size_t)((g.n0 + 31) / 32) * ((g.n1 + 31) / 32);
    return tiles * pair_tab_floats(tile_ref_kind(v)) * sizeof(float);
}

hipError_t launch_pair_table(int v, const Geom& g, float* tab, hipStream_t st)
{
    if (pair_table_bytes(v, g) == 0) return hipErrorInvalidValue;
    const bool a = false;  // the table does not depend on the amplitudes
    IDISIM_DISPATCH_PAIR(launch_pair_table_q, g, tab, st)
    return hipErrorInvalidValue;
}

hipError_t launch_batched_pair(int v, bool a, const Geom& g, const double* emit, const double* amp,
                               const WorkItem* items, int nitems, int* counter, int nblocks, void* out, int kind,
                               hipStream_t st)
{
    if (g.foldw == 0 && (g.nf || g.secondorder)) return hipErrorInvalidValue;
    if (v == kTileBlock)
        return a ? launch_batched_block_a<true>(g, emit, amp, items, nitems, counter, nblocks, out, kind, st)
                 : launch_batched_block_a<false>(g, emit, amp, items, nitems, counter, nblocks, out, kind, st);
    IDISIM_DISPATCH_PAIR(launch_batched_pair_q, g, emit, amp, items, nitems, counter, nblocks, out, kind, st)
    return hipErrorInvalidValue;
}

int occupancy_batched_pair(int v, bool a)
{
    if (v == kTileBlock) return a ? occupancy_block_a<true>() : occupancy_block_a<false>();
    IDISIM_DISPATCH_PAIR(occupancy_pair_q)
    return 1;
}

hipError_t launch_finalize_split(const Geom& g, const float4* partial, int splits, void* out, int kind,
                                 hipStream_t st)
{
    const size_t npix = (size_t)g.n0 * g.n1;
    hipLaunchKernelGGL(k_finalize_split, dim3((unsigned)((npix + 63) / 64)), dim3(256), 0, st, g, partial, splits,
                       out, kind);
    return hipGetLastError();
}

template <int V>
static hipError_t launch_fp64_v(const Geom& g, const double* emit, const double* amp, int64_t n, int splits,
                                double4* partial, void* out, int kind, hipStream_t st)
{
    constexpr int PX = tile_px(V), PY = tile_py(V);
    const int tiles_x = (g.n0 + 16 * PX - 1) / (16 * PX);
    const int tiles_y = (g.n1 + 16 * PY - 1) / (16 * PY);
    const int64_t per = split_size(n, splits);
    dim3 grid(tiles_x * tiles_y, splits);
    hipLaunchKernelGGL((k_fwd_fp64<PX, PY>), grid, dim3(kBlock), 0, st, g, emit, amp, n, tiles_y,
                       per, partial, out, kind);
    return hipGetLastError();
}

template <int V>
static hipError_t launch_batched_fp64_v(const Geom& g, const double* emit, const double* amp,
                                        const WorkItem* items, int nitems, int* counter, int nblocks, void* out,
                                        int kind, hipStream_t st)
{
    constexpr int PX = tile_px(V), PY = tile_py(V);
    const int tiles_y = (g.n1 + 16 * PY - 1) / (16 * PY);
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, st, counter);
    hipLaunchKernelGGL((k_batched_fp64<PX, PY>), dim3(nblocks), dim3(kBlock), 0, st, g, emit, amp, items,
                       nitems, counter, tiles_y, out, kind);
    return hipGetLastError();
}

template <int V>
static int occupancy_fp64_v()
{
    constexpr int PX = tile_px(V), PY = tile_py(V);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batched_fp64<PX, PY>, kBlock, 0) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

hipError_t launch_fwd_fp64(int v, const Geom& g, const double* emit, const double* amp, int64_t n, int splits,
                           double4* partial, void* out, int kind, hipStream_t st)
{
    switch (v) {
        case 0: return launch_fp64_v<0>(g, emit, amp, n, splits, partial, out, kind, st);
        case 1: return launch_fp64_v<1>(g, emit, amp, n, splits, partial, out, kind, st);
        case 2: return launch_fp64_v<2>(g, emit, amp, n, splits, partial, out, kind, st);
        case 3: return launch_fp64_v<3>(g, emit, amp, n, splits, partial, out, kind, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_batched_fp64(int v, const Geom& g, const double* emit, const double* amp,
                               const WorkItem* items, int nitems, int* counter, int nblocks, void* out, int kind,
                               hipStream_t st)
{
    switch (v) {
        case 0: return launch_batched_fp64_v<0>(g, emit, amp, items, nitems, counter, nblocks, out, kind, st);
        case 1: return launch_batched_fp64_v<1>(g, emit, amp, items, nitems, counter, nblocks, out, kind, st);
        case 2: return launch_batched_fp64_v<2>(g, emit, amp, items, nitems, counter, nblocks, out, kind, st);
        case 3: return launch_batched_fp64_v<3>(g, emit, amp, items, nitems, counter, nblocks, out, kind, st);
        default: return hipErrorInvalidValue;
    }
}

int occupancy_batched_fp64(int v)
{
    switch (v) {
        case 0: return occupancy_fp64_v<0>();
        case 1: return occupancy_fp64_v<1>();
        case 2: return occupancy_fp64_v<2>();
        case 3: return occupancy_fp64_v<3>();
        default: return 1;
    }
}

hipError_t launch_finalize_fp64(const Geom& g, const double4* partial, int splits, void* out, int kind,
                                hipStream_t st)
{
    const size_t npix = (size_t)g.n0 * g.n1;
    hipLaunchKernelGGL(k_finalize_fp64, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, st, g, partial,
                       splits, out, kind);
    return hipGetLastError();
}

hipError_t launch_extent(const double* emit, int64_t n, unsigned long long* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(1), 0, st, out, 0ull);
    int64_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > 1024) blocks = 1024;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_extent, dim3((unsigned)blocks), dim3(kBlock), 0, st, emit, n, out);
    return hipGetLastError();
}

hipError_t launch_zero(void* out, size_t bytes, hipStream_t st)
{
    // every output size here is a multiple of 4 bytes (float / double / complex elements)
    const int64_t words = (int64_t)(bytes / 4);
    if (words == 0) return hipSuccess;
    int64_t blocks = (words + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_zero_words, dim3((unsigned)blocks), dim3(256), 0, st, static_cast<uint32_t*>(out), words);
    return hipGetLastError();
}

}  // namespace idisim
