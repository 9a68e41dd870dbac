// Adjoint of I = |E|^2 (no reference counterpart): for a loss L(I) with
// G_p = dL/dI_p and w_p = G_p conj(E_p),
//   dL/dphi_j = -2 a_j sum_p Im(w_p e^{i theta_pj})
//   dL/dr_j   = -2 a_j sum_p Im(w_p e^{i theta_pj}) grad_r theta_pj
//   dL/da_j   =  2     sum_p Re(w_p e^{i theta_pj})
// grad_r theta_pj = k (z_hat - (r_p - r_j)/D_pj)  (near field)
//                 = -q_p (+ k/|r_p| (x_j, y_j, 0) with secondorder)  (far field).
// With the tile decomposition of geometry.hpp (u_p = u_c + du_p, dq_p = -krev du_p):
//   near field:  z_hat - n_pj = -u_c + beta_cj - du_p,
//                beta_cj = (rh_c (w_cj - s_cj) + r_j) / D_cj   (O(|r_j|/R), taken at the centre)
//   far field:   -u_p = -u_c - du_p
// so per (tile, emitter) only S = sum_p w_p e^{i theta} (complex) and
// V = sum_p Im(.) dq_p (3 floats) are needed in FP32; they are folded into
// FP64 accumulators with the FP64 tile-centre factors after every tile.
//
// Layout: emitter-parallel.  A workgroup owns 256*EPL emitters (EPL per lane)
// and a range of 16 x 16 pixel tiles; each tile's per-pixel data (dq, lam, w')
// is formed in FP64 by one lane per pixel and staged in LDS, then read by all
// lanes as broadcasts.  Partial gradients of each pixel range go to a scratch
// buffer [splits][n][5] and a finalize kernel sums them in FP64.
#include "kernels.hpp"

namespace idisim {

template <int EPL, bool HAS_W>
__global__ __launch_bounds__(kBlock) void k_adjoint(Geom g, const double* __restrict__ emit,
                                                    const double* __restrict__ amp, int64_t n,
                                                    const float2* __restrict__ field,
                                                    const float* __restrict__ dLdI, int tiles_y, int ntiles,
                                                    int tiles_per_split, double* __restrict__ partial)
{
    __shared__ float4 sp0[kBlock];  // dqx, dqy, dqz, lam
    __shared__ float2 sp1[kBlock];  // w'_p (per-pixel phase removed)
    const int tid = threadIdx.x;
    const int64_t ebase = (int64_t)blockIdx.x * (kBlock * EPL);

    double gx[EPL], gy[EPL], gz[EPL], gph[EPL], gam[EPL];
#pragma unroll
    for (int i = 0; i < EPL; ++i) gx[i] = gy[i] = gz[i] = gph[i] = gam[i] = 0.0;

    const int t0 = blockIdx.y * tiles_per_split;
    const int t1 = (t0 + tiles_per_split < ntiles) ? t0 + tiles_per_split : ntiles;
    for (int t = t0; t < t1; ++t) {
        const int tx0 = (t / tiles_y) * 16, ty0 = (t % tiles_y) * 16;
        const int xe = min(tx0 + 16, g.n0) - 1, ye = min(ty0 + 16, g.n1) - 1;
        const Frame c = pixel_frame(g, 0.5 * (tx0 + xe), 0.5 * (ty0 + ye));

        __syncthreads();  // previous tile consumed
        {
            const int x = tx0 + (tid >> 4), y = ty0 + (tid & 15);
            float4 q = make_float4(0.f, 0.f, 0.f, 1.f);
            float2 w = make_float2(0.f, 0.f);
            if (x < g.n0 && y < g.n1) {
                const Frame f = pixel_frame(g, (double)x, (double)y);
                q = make_float4((float)(-g.krev * (f.u[0] - c.u[0])), (float)(-g.krev * (f.u[1] - c.u[1])),
                                (float)(-g.krev * (f.u[2] - c.u[2])), (float)(c.R / f.R));
                const size_t idx = (size_t)x * g.n1 + y;
                const float2 E = field[idx];
                const double G = (double)dLdI[idx];
                double wr = G * (double)E.x, wi = -G * (double)E.y;
                if (g.nf) {
                    double ph = g.krev * f.RmZ;
                    ph -= rint(ph);
                    double sn, cs;
                    sincos(kTwoPi * ph, &sn, &cs);
                    const double r2 = wr * cs - wi * sn;
                    wi = wr * sn + wi * cs;
                    wr = r2;
                }
                w = make_float2((float)wr, (float)wi);
            }
            sp0[tid] = q;
            sp1[tid] = w;
        }

        float rx[EPL], ry[EPL], rz[EPL], rA[EPL], rW[EPL];
#pragma unroll
        for (int i = 0; i < EPL; ++i) {
            const int64_t e = ebase + tid + (int64_t)i * kBlock;
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
        __syncthreads();

        float Sre[EPL], Sim[EPL], Vx[EPL], Vy[EPL], Vz[EPL], Lm[EPL];
#pragma unroll
        for (int i = 0; i < EPL; ++i) Sre[i] = Sim[i] = Vx[i] = Vy[i] = Vz[i] = Lm[i] = 0.f;

#pragma unroll 2
        for (int p = 0; p < kBlock; ++p) {
            const float4 q = sp0[p];
            const float2 w = sp1[p];
#pragma unroll
            for (int i = 0; i < EPL; ++i) {
                float tt = HAS_W ? fmaf(q.w, rW[i], rA[i]) : rA[i];
                tt = fmaf(q.x, rx[i], tt);
                tt = fmaf(q.y, ry[i], tt);
                tt = fmaf(q.z, rz[i], tt);
                const float sn = __builtin_amdgcn_sinf(tt);
                const float cs = __builtin_amdgcn_cosf(tt);
                const float re = fmaf(w.x, cs, -w.y * sn);
                const float im = fmaf(w.x, sn, w.y * cs);
                Sre[i] += re;
                Sim[i] += im;
                Vx[i] = fmaf(im, q.x, Vx[i]);
                Vy[i] = fmaf(im, q.y, Vy[i]);
                Vz[i] = fmaf(im, q.z, Vz[i]);
                if (HAS_W && !g.nf) Lm[i] = fmaf(im, q.w, Lm[i]);
            }
        }

        // fold the tile into FP64 accumulators
#pragma unroll
        for (int i = 0; i < EPL; ++i) {
            const int64_t e = ebase + tid + (int64_t)i * kBlock;
            if (e >= n) continue;
            const double* a4 = emit + 4 * e;
            const double ex = a4[0], ey = a4[1], ez = a4[2];
            double bx = 0.0, by = 0.0, bz = 0.0;
            if (g.nf) {
                const double ucr = c.u[0] * ex + c.u[1] * ey + c.u[2] * ez;
                const double s = ucr + ez;
                const double hx = c.rh[1] * ez - c.rh[2] * ey;
                const double hy = c.rh[2] * ex - c.rh[0] * ez;
                const double hz = c.rh[0] * ey - c.rh[1] * ex;
                const double h2 = hx * hx + hy * hy + hz * hz;
                const double a = c.R - s;
                const double D = sqrt(a * a + h2);
                const double w = (a >= 0.0) ? h2 / (D + a) : D - a;
                const double iD = 1.0 / D;
                bx = (c.rh[0] * (w - s) + ex) * iD;
                by = (c.rh[1] * (w - s) + ey) * iD;
                bz = (c.rh[2] * (w - s) + ez) * iD;
            }
            const double S_im = (double)Sim[i];
            const double ik = 1.0 / g.krev;
            // sum_p Im_p grad theta_pj = k [(-u_c + beta) S_im + V / krev]
            gx[i] += g.k * ((bx - c.u[0]) * S_im + (double)Vx[i] * ik);
            gy[i] += g.k * ((by - c.u[1]) * S_im + (double)Vy[i] * ik);
            gz[i] += g.k * ((bz - c.u[2]) * S_im + (double)Vz[i] * ik);
            if (HAS_W && !g.nf) {
                // secondorder: theta += k/(2 R_p)(x^2+y^2), R_c/R_p = lam
                const double L = (double)Lm[i] * g.k / c.R;
                gx[i] += L * ex;
                gy[i] += L * ey;
            }
            gph[i] += S_im;
            gam[i] += (double)Sre[i];
        }
    }

#pragma unroll
    for (int i = 0; i < EPL; ++i) {
        const int64_t e = ebase + tid + (int64_t)i * kBlock;
        if (e >= n) continue;
        const double a = (amp != nullptr) ? amp[e] : 1.0;
        double* dst = partial + ((size_t)blockIdx.y * n + e) * 5;
        dst[0] = -2.0 * a * gx[i];
        dst[1] = -2.0 * a * gy[i];
        dst[2] = -2.0 * a * gz[i];
        dst[3] = -2.0 * a * gph[i];
        dst[4] = 2.0 * gam[i];
    }
}

__global__ void k_adjoint_finalize(const double* __restrict__ partial, int splits, int64_t n,
                                   float* __restrict__ grad_pos, float* __restrict__ grad_phase,
                                   float* __restrict__ grad_amp)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double v[5] = {0, 0, 0, 0, 0};
    for (int s = 0; s < splits; ++s) {
        const double* src = partial + ((size_t)s * n + e) * 5;
#pragma unroll
        for (int c = 0; c < 5; ++c) v[c] += src[c];
    }
    grad_pos[3 * e] = (float)v[0];
    grad_pos[3 * e + 1] = (float)v[1];
    grad_pos[3 * e + 2] = (float)v[2];
    grad_phase[e] = (float)v[3];
    if (grad_amp != nullptr) grad_amp[e] = (float)v[4];
}

constexpr int kAdjEPL = 2;

hipError_t launch_adjoint(bool w, const Geom& g, const double* emit, const double* amp, int64_t n,
                          const float2* field, const float* dLdI, int pixel_splits, double* partial,
                          float* grad_pos, float* grad_phase, float* grad_amp, hipStream_t st)
{
    const int tiles_x = (g.n0 + 15) / 16, tiles_y = (g.n1 + 15) / 16;
    const int ntiles = tiles_x * tiles_y;
    if (pixel_splits < 1) pixel_splits = 1;
    if (pixel_splits > ntiles) pixel_splits = ntiles;
    const int per = (ntiles + pixel_splits - 1) / pixel_splits;
    const int splits = (ntiles + per - 1) / per;
    dim3 grid((unsigned)((n + kBlock * kAdjEPL - 1) / (kBlock * kAdjEPL)), (unsigned)splits);
    if (w)
        hipLaunchKernelGGL((k_adjoint<kAdjEPL, true>), grid, dim3(kBlock), 0, st, g, emit, amp, n, field, dLdI,
                           tiles_y, ntiles, per, partial);
    else
        hipLaunchKernelGGL((k_adjoint<kAdjEPL, false>), grid, dim3(kBlock), 0, st, g, emit, amp, n, field, dLdI,
                           tiles_y, ntiles, per, partial);
    hipLaunchKernelGGL(k_adjoint_finalize, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, partial, splits,
                       n, grad_pos, grad_phase, grad_amp);
    return hipGetLastError();
}

}  // namespace idisim
