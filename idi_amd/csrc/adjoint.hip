// Adjoint of I = |E|^2 (no reference counterpart): for a loss L(I) with
// G_p = dL/dI_p and w_p = G_p conj(E_p),
//   dL/dphi_j = -2 a_j sum_p Im(w_p e^{i theta_pj})
//   dL/dr_j   = -2 a_j sum_p Im(w_p e^{i theta_pj}) grad_r theta_pj
//   dL/da_j   =  2     sum_p Re(w_p e^{i theta_pj})
// grad_r theta_pj = k (z_hat - (r_p - r_j)/D_pj)  (near field)
//                 = -q_p (+ k/|r_p| (x_j, y_j, 0) with secondorder)  (far field).
// With the tile decomposition of geometry.hpp (u_p = u_c + du_p, dq_p = -krev du_p):
//   near field:  z_hat - n_pj = -u_c + beta_cj - du_p,
//                beta_cj = (rh_c (w_cj - s_cj) + r_j) / D_cj = r_j / D_cj - rh_c sigma_cj
//                (O(|r_j|/R), taken at the tile centre)
//   far field:   -u_p = -u_c - du_p
// so per (tile, emitter) only S = sum_p Im(w_p e^{i theta}) and V = sum_p Im(.) dq_p
// (3 floats) are needed in FP32 (plus sum Re(.) for dL/da); they are folded into FP64
// accumulators with the FP64 tile-centre vector u_c after every tile.
// With the 1/r "scale" model the amplitude is R_c/(D_cj R_p): 1/R_p goes into w'_p and R_c/D_cj
// multiplies the tile's sums when they are folded.  The gradient of the amplitude itself,
// Re(.) grad(1/D), is left out: next to the phase term it is smaller by 1/(k D) (about 1e-9
// for X-rays at a detector distance of 0.1 m).
// With w'_p = m_p e^{2 pi i psi_p}, Im(w'_p e^{2 pi i tau}) = m_p sin(2 pi (tau + psi_p)):
// one transcendental per pair, and m_p, m_p dq_p are per-pixel table entries, so
// S and V are four FMAs on that sine (the cosine is needed for dL/da only).
//
// Kernels:
//   k_adj_pixels  once per call: per pixel of each 16 x 16 tile, dq_p and lam_p = R_c/R_p
//                 relative to the tile centre and w'_p = w_p e^{2 pi i Phi_p} (the per-pixel
//                 near-field phase, FP64) as (dq, psi), m (dq, 1), (lam, m lam), laid out
//                 tile-major for LDS staging; for the squared-intensity loss it forms G_p from
//                 the field and the target itself, and each tile's part of the loss.
//   k_adjoint     emitter-parallel: a workgroup owns 64*EPL emitters, shared by its four
//                 waves; each wave streams its own pixel tiles (tile t0 + 4 it + wave) through
//                 a private LDS slot, so the 4 waves cover 4 tiles per step.  The per-wave
//                 FP64 sums are reduced through LDS and written as partials [splits][n][5].
//   k_adjoint_finalize  sums the partials over splits in FP64.
// Four emitters per lane measured faster than two or eight at 4096 emitters on 256^2 pixels.
#include "device_util.hpp"

namespace idisim {

namespace {
constexpr int kAdjWaves = kBlock / 64;  // waves per workgroup
constexpr int kAdjEPL = 4;              // emitters per lane
constexpr int kAdjTile = 256;           // pixels per 16 x 16 tile

__device__ inline Frame adj_tile_centre(const Geom& g, int t, int tiles_y)
{
    const int tx0 = (t / tiles_y) * 16, ty0 = (t % tiles_y) * 16;
    const int xe = min(tx0 + 16, g.n0) - 1, ye = min(ty0 + 16, g.n1) - 1;
    return pixel_frame(g, 0.5 * (tx0 + xe), 0.5 * (ty0 + ye));
}
}  // namespace

// The pixel-table entries of pixel (x, y) of 16 x 16 tile t, from its complex64 field value E,
// its frame f and (nf) its pixel_phasor (cs, sn).
// target == nullptr: G_p = dLdI[p].  Otherwise the squared-intensity loss L = sum_p (I_p - T_p)^2:
// G_p = 2 (I_p - T_p) with I_p = |E_p|^2 in FP64, and res2 = (I_p - T_p)^2.
__device__ inline void adj_pixel_terms(const Geom& g, int t, int tiles_y, int x, int y, const Frame& f, double cs,
                                       double sn, float2 E, const float* __restrict__ dLdI,
                                       const float* __restrict__ target, float4& a, float4& b, float2& c2,
                                       double& res2)
{
    const Frame c = adj_tile_centre(g, t, tiles_y);
    const double dqx = -g.krev * (f.u[0] - c.u[0]), dqy = -g.krev * (f.u[1] - c.u[1]);
    const double dqz = -g.krev * (f.u[2] - c.u[2]), lam = c.R / f.R;
    const size_t idx = (size_t)x * g.n1 + y;
    double G;
    if (target != nullptr) {
        const double res = (double)E.x * (double)E.x + (double)E.y * (double)E.y - (double)target[idx];
        res2 = res * res;
        G = 2.0 * res;
    } else {
        G = (double)dLdI[idx];
    }
    double wr = G * (double)E.x, wi = -G * (double)E.y;
    if (g.nf) {
        const double r2 = wr * cs - wi * sn;
        wi = wr * sn + wi * cs;
        wr = r2;
    }
    if (g.scale) {  // the per-pixel factor 1/R_p of the "scale" amplitude R_c/(D_cj R_p)
        wr /= f.R;
        wi /= f.R;
    }
    const double m = sqrt(wr * wr + wi * wi);
    const double psi = (m > 0.0) ? atan2(wi, wr) * kInv2Pi : 0.0;
    a = make_float4((float)dqx, (float)dqy, (float)dqz, (float)psi);
    b = make_float4((float)(m * dqx), (float)(m * dqy), (float)(m * dqz), (float)m);
    c2 = make_float2((float)lam, (float)(m * lam));
}

// Stores the thread's table entries and, with a target, the tile's sum of squared residuals
// (an LDS tree in a fixed order) into loss_tile[t].
__device__ inline void adj_pixel_store(int t, int l, float4 a, float4 b, float2 c2, double res2, bool loss,
                                       float4* __restrict__ pa, float4* __restrict__ pb, float2* __restrict__ pc,
                                       double* __restrict__ loss_tile, double* red)
{
    pa[(size_t)t * kAdjTile + l] = a;
    pb[(size_t)t * kAdjTile + l] = b;
    pc[(size_t)t * kAdjTile + l] = c2;
    if (loss) {
        red[l] = res2;
        __syncthreads();
        for (int s = kAdjTile / 2; s > 0; s >>= 1) {
            if (l < s) red[l] += red[l + s];
            __syncthreads();
        }
        if (l == 0) loss_tile[t] = red[0];
    }
}

// The pixel table from the field of a forward call (one workgroup per 16 x 16 tile).
__global__ __launch_bounds__(kAdjTile) void k_adj_pixels(Geom g, const float2* __restrict__ field,
                                                         const float* __restrict__ dLdI,
                                                         const float* __restrict__ target, int tiles_y,
                                                         float4* __restrict__ pa, float4* __restrict__ pb,
                                                         float2* __restrict__ pc, double* __restrict__ loss_tile)
{
    __shared__ double red[kAdjTile];
    const int t = blockIdx.x, l = threadIdx.x;
    const int x = (t / tiles_y) * 16 + (l >> 4), y = (t % tiles_y) * 16 + (l & 15);
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 b = make_float4(0.f, 0.f, 0.f, 0.f);
    float2 c2 = make_float2(1.f, 0.f);
    double res2 = 0.0;
    if (x < g.n0 && y < g.n1) {
        const Frame f = pixel_frame(g, (double)x, (double)y);
        double cs = 1.0, sn = 0.0;
        if (g.nf) pixel_phasor(g, f, cs, sn);
        adj_pixel_terms(g, t, tiles_y, x, y, f, cs, sn, field[(size_t)x * g.n1 + y], dLdI, target, a, b, c2, res2);
    }
    adj_pixel_store(t, l, a, b, c2, res2, target != nullptr, pa, pb, pc, loss_tile, red);
}

// The same table straight from a forward launch's split partials (float4 [splits][npix], the
// mirror-pair kernels' form): each pixel's field is summed and formed as k_finalize_split and
// split_field do (the same bits), stored when field != nullptr, and used at once.  One launch
// where the forward's finalize and k_adj_pixels took two (5 us each at 256^2).
__global__ __launch_bounds__(kAdjTile) void k_adj_pixels_split(Geom g, const float4* __restrict__ partial,
                                                               int splits, float2* __restrict__ field,
                                                               const float* __restrict__ target, int tiles_y,
                                                               float4* __restrict__ pa, float4* __restrict__ pb,
                                                               float2* __restrict__ pc,
                                                               double* __restrict__ loss_tile)
{
    __shared__ double red[kAdjTile];
    const int t = blockIdx.x, l = threadIdx.x;
    const int x = (t / tiles_y) * 16 + (l >> 4), y = (t % tiles_y) * 16 + (l & 15);
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 b = make_float4(0.f, 0.f, 0.f, 0.f);
    float2 c2 = make_float2(1.f, 0.f);
    double res2 = 0.0;
    if (x < g.n0 && y < g.n1) {
        const size_t npix = (size_t)g.n0 * g.n1, idx = (size_t)x * g.n1 + y;
        double re, im;
        split_sum(partial, splits, npix, idx, re, im);
        const Frame f = pixel_frame(g, (double)x, (double)y);
        double cs = 1.0, sn = 0.0;
        if (g.nf) pixel_phasor(g, f, cs, sn);
        const float2 E = split_field(g, f, cs, sn, re, im);
        if (field != nullptr) field[idx] = E;
        adj_pixel_terms(g, t, tiles_y, x, y, f, cs, sn, E, nullptr, target, a, b, c2, res2);
    }
    adj_pixel_store(t, l, a, b, c2, res2, true, pa, pb, pc, loss_tile, red);
}

// HAS_W: lam_p W_cj enters the phase (nf or secondorder, not folded into A).
// LAM:   secondorder far field unfolded: also accumulate sum_p Im(.) lam_p.
// AMP:   also accumulate sum_p Re(.) for dL/da.
// Held to 128 VGPRs (four waves per SIMD; the unfolded secondorder far field with amplitudes
// would spill, and keeps three).
template <int EPL, bool HAS_W, bool LAM, bool AMP>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu((LAM && AMP) ? 3 : 4)))
void k_adjoint(Geom g, const double* __restrict__ emit, const double* __restrict__ amp, int64_t n,
               const float4* __restrict__ pa, const float4* __restrict__ pb, const float2* __restrict__ pc,
               int tiles_y, int ntiles,
                                                    int tiles_per_split, double* __restrict__ partial)
{
    constexpr int M = 64 * EPL;  // emitters per workgroup
    constexpr bool USE_LAM = HAS_W || LAM;
    // the cross-wave reduction after the tile loop reuses the staging area of sa
    union StageOrReduce {
        float4 a[kAdjWaves][kAdjTile];
        double red[kAdjWaves][M];
    };
    __shared__ StageOrReduce u;
    __shared__ float4 sb[kAdjWaves][kAdjTile];
    __shared__ float2 sc[kAdjWaves][USE_LAM ? kAdjTile : 1];
    float4(&sa)[kAdjWaves][kAdjTile] = u.a;
    double(&red)[kAdjWaves][M] = u.red;

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t ebase = (int64_t)blockIdx.x * M;
    __builtin_amdgcn_s_setprio(3);  // the records: see prio_by_progress

    double acc[5][EPL];
#pragma unroll
    for (int i = 0; i < EPL; ++i) acc[0][i] = acc[1][i] = acc[2][i] = acc[3][i] = acc[4][i] = 0.0;

    const int t0 = blockIdx.y * tiles_per_split;
    const int t1 = min(t0 + tiles_per_split, ntiles);
    const int iters = (t1 - t0 + kAdjWaves - 1) / kAdjWaves;  // the same for every wave
    for (int it = 0; it < iters; ++it) {
        const int t = t0 + it * kAdjWaves + wave;
        const bool active = t < t1;  // wave-uniform
        __syncthreads();             // this wave's previous tile consumed (uniform barrier count)
        if (active) {
#pragma unroll
            for (int k = 0; k < kAdjTile / 64; ++k) {
                const int l = lane + 64 * k;
                sa[wave][l] = pa[(size_t)t * kAdjTile + l];
                sb[wave][l] = pb[(size_t)t * kAdjTile + l];
                if (USE_LAM) sc[wave][l] = pc[(size_t)t * kAdjTile + l];
            }
        }
        __syncthreads();
        if (!active) continue;

        float rx[EPL], ry[EPL], rz[EPL], rA[EPL], rW[EPL];
        {
            const Frame c = adj_tile_centre(g, t, tiles_y);
#pragma unroll
            for (int i = 0; i < EPL; ++i) {
                const int64_t e = ebase + lane + 64 * i;
                if (e < n) {
                    const double* a4 = emit + 4 * e;
                    const SplitRec r = make_split_rec(g, c, a4[0], a4[1], a4[2], a4[3], 1.0);
                    rx[i] = r.x;
                    ry[i] = r.y;
                    rz[i] = r.z;
                    rA[i] = r.A;
                    rW[i] = r.W;
                } else {
                    rx[i] = ry[i] = rz[i] = rA[i] = rW[i] = 0.f;
                }
            }
        }

        float Sim[EPL], Vx[EPL], Vy[EPL], Vz[EPL], Sre[EPL], Lm[EPL];
#pragma unroll
        for (int i = 0; i < EPL; ++i) Sim[i] = Vx[i] = Vy[i] = Vz[i] = Sre[i] = Lm[i] = 0.f;

        const float4* a4p = sa[wave];
        const float4* b4p = sb[wave];
        // the wave's priority falls as its pixel loops progress (blocks of 8 pixels), so that the
        // waves sharing a SIMD finish together (measured 79 -> 74 us at 4096 emitters on 256^2)
        for (int p0 = 0; p0 < kAdjTile; p0 += 8) {
            prio_by_progress((32 * (it * kAdjTile + p0)) / (iters * kAdjTile));
        // two pixels per iteration: the second pixel's loads and phases overlap the first's sines
        // (measured 88 -> 82 us per 4096-emitter 256^2 launch against no unrolling)
#pragma unroll 2
        for (int p = p0; p < p0 + 8; ++p) {
            const float4 a = a4p[p];  // dq, psi
            const float4 b = b4p[p];  // m dq, m
            float2 c2 = make_float2(1.f, 0.f);
            if (USE_LAM) c2 = sc[wave][p];  // lam, m lam
#pragma unroll
            for (int i = 0; i < EPL; ++i) {
                float tt = rA[i] + a.w;
                if (HAS_W) tt = fmaf(c2.x, rW[i], tt);
                tt = fmaf(a.x, rx[i], tt);
                tt = fmaf(a.y, ry[i], tt);
                tt = fmaf(a.z, rz[i], tt);
                const float sn = __builtin_amdgcn_sinf(tt);
                Sim[i] = fmaf(b.w, sn, Sim[i]);
                Vx[i] = fmaf(b.x, sn, Vx[i]);
                Vy[i] = fmaf(b.y, sn, Vy[i]);
                Vz[i] = fmaf(b.z, sn, Vz[i]);
                if (LAM) Lm[i] = fmaf(c2.y, sn, Lm[i]);
                if (AMP) Sre[i] = fmaf(b.w, __builtin_amdgcn_cosf(tt), Sre[i]);
            }
        }
        }
        // the fold and the next tile's records (FP64, latency-bound) at the top priority
        __builtin_amdgcn_s_setprio(3);

        // fold the tile into the FP64 accumulators:
        // sum_p Im_p grad theta_pj = k [(beta - u_c) S + V / krev]
        // The tile centre and the records are formed again, from a tile index the compiler cannot
        // equate with t (an empty asm): the values the fold needs (beta, the amplitude factor, u_c)
        // are then not held in VGPRs across the pixel loop (one FP64 record per emitter and tile,
        // against 256 pairs).
        int tf = t;
        asm volatile("" : "+v"(tf));
        const Frame c = adj_tile_centre(g, tf, tiles_y);
        const double ik = 1.0 / g.krev;
#pragma unroll
        for (int i = 0; i < EPL; ++i) {
            const int64_t e = ebase + lane + 64 * i;
            float bx = 0.f, by = 0.f, bz = 0.f, rm = 0.f;
            if (e < n) {
                const double* a4 = emit + 4 * e;
                const SplitRec r = make_split_rec(g, c, a4[0], a4[1], a4[2], a4[3], 1.0);
                rm = r.amp;  // R_c / D_cj with "scale", else 1
                // beta_cj in FP32 (|beta| ~ |r_j|/R; only its relative precision matters)
                const float fx = (float)a4[0], fy = (float)a4[1], fz = (float)a4[2];
                bx = fx * r.iD - (float)c.rh[0] * r.sigma;
                by = fy * r.iD - (float)c.rh[1] * r.sigma;
                bz = fz * r.iD - (float)c.rh[2] * r.sigma;
            }
            const double ma = (double)rm;
            const double S = ma * (double)Sim[i];
            const double Vfx = ma * (double)Vx[i], Vfy = ma * (double)Vy[i], Vfz = ma * (double)Vz[i];
            acc[0][i] += g.k * (((double)bx - c.u[0]) * S + Vfx * ik);
            acc[1][i] += g.k * (((double)by - c.u[1]) * S + Vfy * ik);
            acc[2][i] += g.k * (((double)bz - c.u[2]) * S + Vfz * ik);
            if (!g.nf && g.secondorder) {
                // theta += k/(2 R_p)(x^2 + y^2) with R_c/R_p = lam (lam ~ 1 when W is folded)
                const double L = (LAM ? ma * (double)Lm[i] : S) * g.k / c.R;
                acc[0][i] += L * (double)rx[i];
                acc[1][i] += L * (double)ry[i];
            }
            acc[3][i] += S;
            if (AMP) acc[4][i] += ma * (double)Sre[i];
        }
    }

    // reduce the four waves' sums (same emitters) and write this split's partials; the
    // first barrier also ends every wave's use of sa
#pragma unroll
    for (int comp = 0; comp < 5; ++comp) {
        if (comp == 4 && !AMP) break;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < EPL; ++i) red[wave][lane + 64 * i] = acc[comp][i];
        __syncthreads();
        for (int m = tid; m < M; m += kBlock) {
            const int64_t e = ebase + m;
            if (e >= n) continue;
            double v = 0.0;
#pragma unroll
            for (int wv = 0; wv < kAdjWaves; ++wv) v += red[wv][m];
            const double a = (amp != nullptr) ? amp[e] : 1.0;
            partial[((size_t)blockIdx.y * n + e) * 5 + comp] = (comp == 4) ? 2.0 * v : -2.0 * a * v;
        }
    }
}

// Sum of the partials over splits: one column per (emitter, component) value, four rows of
// threads stride over the splits so that enough loads are in flight.  With loss != nullptr,
// workgroup 0 also sums the tiles' squared residuals into *loss, in a fixed order.
__global__ __launch_bounds__(256) void k_adjoint_finalize(const double* __restrict__ partial, int splits,
                                                          int64_t n, float* __restrict__ grad_pos,
                                                          float* __restrict__ grad_phase,
                                                          float* __restrict__ grad_amp,
                                                          const double* __restrict__ loss_tile, int ntiles,
                                                          double* __restrict__ loss)
{
    __shared__ double red[4][64];
    if (loss != nullptr && blockIdx.x == 0) {
        double v = 0.0;
        for (int t = threadIdx.x; t < ntiles; t += 256) v += loss_tile[t];
        double* lr = &red[0][0];
        lr[threadIdx.x] = v;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) lr[threadIdx.x] += lr[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) *loss = lr[0];
        __syncthreads();
    }
    const int col = threadIdx.x & 63, row = threadIdx.x >> 6;
    const int64_t total = n * 5;
    const int64_t f = (int64_t)blockIdx.x * 64 + col;  // e * 5 + component
    double v0 = 0.0, v1 = 0.0;
    if (f < total) {
        const double* src = partial + f;
        int s = row;
        for (; s + 4 < splits; s += 8) {
            v0 += src[(size_t)s * total];
            v1 += src[(size_t)(s + 4) * total];
        }
        if (s < splits) v0 += src[(size_t)s * total];
    }
    red[row][col] = v0 + v1;
    __syncthreads();
    if (row == 0 && f < total) {
        const double v = (red[0][col] + red[1][col]) + (red[2][col] + red[3][col]);
        const int64_t e = f / 5;
        const int c = (int)(f - 5 * e);
        if (c < 3)
            grad_pos[3 * e + c] = (float)v;
        else if (c == 3)
            grad_phase[e] = (float)v;
        else if (grad_amp != nullptr)
            grad_amp[e] = (float)v;
    }
}

int adjoint_occupancy(bool w, bool lam, bool amp)
{
    static int cache[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int key = (w ? 4 : 0) | (lam ? 2 : 0) | (amp ? 1 : 0);
    if (cache[key] > 0) return cache[key];
    int nb = 0;
    hipError_t e;
    if (!w)
        e = amp ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_adjoint<kAdjEPL, false, false, true>, kBlock, 0)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_adjoint<kAdjEPL, false, false, false>, kBlock, 0);
    else if (!lam)
        e = amp ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_adjoint<kAdjEPL, true, false, true>, kBlock, 0)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_adjoint<kAdjEPL, true, false, false>, kBlock, 0);
    else
        e = amp ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_adjoint<kAdjEPL, true, true, true>, kBlock, 0)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_adjoint<kAdjEPL, true, true, false>, kBlock, 0);
    if (e != hipSuccess || nb < 1) nb = 1;
    cache[key] = nb;
    return nb;
}

AdjointLayout adjoint_layout(const Geom& g, int64_t n, int num_cus, int blocks_per_cu)
{
    AdjointLayout L;
    L.tiles_y = (g.n1 + 15) / 16;
    L.ntiles = ((g.n0 + 15) / 16) * L.tiles_y;
    L.egroups = (n + 64 * kAdjEPL - 1) / (64 * kAdjEPL);
    // Each workgroup takes kAdjWaves * it tiles (it per wave).  Pick it to minimise the
    // makespan rounds * it on num_cus * blocks_per_cu resident workgroups; on ties, fewer
    // splits (less partial traffic).
    const int64_t slots = (int64_t)num_cus * (blocks_per_cu > 0 ? blocks_per_cu : 1);
    const int max_it = (L.ntiles + kAdjWaves - 1) / kAdjWaves;
    int best_it = max_it;
    int64_t best = INT64_MAX;
    for (int it = 1; it <= max_it; ++it) {
        const int64_t splits = (L.ntiles + kAdjWaves * it - 1) / (kAdjWaves * it);
        if (it > 1 && splits == (L.ntiles + kAdjWaves * (it - 1) - 1) / (kAdjWaves * (it - 1))) continue;
        const int64_t rounds = (L.egroups * splits + slots - 1) / slots;
        const int64_t span = rounds * it;
        if (span <= best) {
            best = span;
            best_it = it;
        }
    }
    L.per = kAdjWaves * best_it;
    L.splits = (L.ntiles + L.per - 1) / L.per;
    L.pix_bytes = (size_t)L.ntiles * kAdjTile * (2 * sizeof(float4) + sizeof(float2));
    L.part_bytes = (size_t)L.splits * (size_t)n * 5 * sizeof(double);
    L.bytes = L.pix_bytes + L.part_bytes + (size_t)L.ntiles * sizeof(double);  // + the tiles' loss parts
    return L;
}

#define IDISIM_ADJ_LAUNCH(W, LM, A)                                                                        \
    hipLaunchKernelGGL((k_adjoint<kAdjEPL, W, LM, A>), grid, dim3(kBlock), 0, st, g, emit, amp, n, pa, pb, pc, \
                       L.tiles_y, L.ntiles, L.per, partial)

hipError_t launch_adjoint(bool w, bool lam, const Geom& g, const AdjointLayout& L, const double* emit,
                          const double* amp, int64_t n, const float2* field, const float* dLdI, const float* target,
                          double* loss, void* scratch, float* grad_pos, float* grad_phase, float* grad_amp,
                          hipStream_t st, const float4* fwd_partial, int fwd_splits, float2* field_out)
{
    if (fwd_partial != nullptr && target == nullptr) return hipErrorInvalidValue;
    const size_t np = (size_t)L.ntiles * kAdjTile;
    float4* pa = static_cast<float4*>(scratch);
    float4* pb = pa + np;
    float2* pc = reinterpret_cast<float2*>(pb + np);
    double* partial = reinterpret_cast<double*>(static_cast<char*>(scratch) + L.pix_bytes);
    double* loss_tile = reinterpret_cast<double*>(static_cast<char*>(scratch) + L.pix_bytes + L.part_bytes);
    if (fwd_partial != nullptr)
        hipLaunchKernelGGL(k_adj_pixels_split, dim3((unsigned)L.ntiles), dim3(kAdjTile), 0, st, g, fwd_partial,
                           fwd_splits, field_out, target, L.tiles_y, pa, pb, pc, loss_tile);
    else
        hipLaunchKernelGGL(k_adj_pixels, dim3((unsigned)L.ntiles), dim3(kAdjTile), 0, st, g, field, dLdI, target,
                           L.tiles_y, pa, pb, pc, loss_tile);
    const dim3 grid((unsigned)L.egroups, (unsigned)L.splits);
    const bool a = grad_amp != nullptr;
    if (!w) {
        if (a) IDISIM_ADJ_LAUNCH(false, false, true);
        else IDISIM_ADJ_LAUNCH(false, false, false);
    } else if (!lam) {
        if (a) IDISIM_ADJ_LAUNCH(true, false, true);
        else IDISIM_ADJ_LAUNCH(true, false, false);
    } else {
        if (a) IDISIM_ADJ_LAUNCH(true, true, true);
        else IDISIM_ADJ_LAUNCH(true, true, false);
    }
    hipLaunchKernelGGL(k_adjoint_finalize, dim3((unsigned)((5 * n + 63) / 64)), dim3(256), 0, st, partial, L.splits,
                       n, grad_pos, grad_phase, grad_amp, loss_tile, L.ntiles, target != nullptr ? loss : nullptr);
    return hipGetLastError();
}

#undef IDISIM_ADJ_LAUNCH

}  // namespace idisim
