// FP32-split forward kernels with the phase contraction on the matrix cores (gfx950 MFMA).
//
// With the tile decomposition of geometry.hpp the per-pair phase is a rank-4
// contraction plus a per-emitter constant:
//     tau_pj = [dq_p, lam_p] . [r'_j, W_j] + A_j
// so one v_mfma_f32_16x16x4_f32 forms tau for 16 pixels x 16 emitters:
//   A operand (16 x 4): pixel m, k = 0..3 -> (dq_x, dq_y, dq_z, lam) ; lane m + 16k
//   B operand (4 x 16): k, emitter n      -> (x', y', z', W)          ; lane n + 16k
//   C operand          : A_n broadcast over the 16 pixels            ; lane n + 16k
//   D (16 x 16)        : lane l holds tau for pixels 4 (l/16) + i, i = 0..3, emitter l % 16.
// The VALU then only evaluates v_sin_f32 / v_cos_f32 and two FMAs (amplitude)
// per pair.  Each lane accumulates 1/16 of the emitters of its pixels; the 16
// partial sums are combined with lane shuffles (in FP64) once per tile.
//
// Workgroup: NW waves; wave w owns R pixel rows of the (R NW) x 16 tile, one
// MFMA per row and block of 16 emitters.  Emitter records are formed in FP64
// by one lane per emitter (chunks of 64 NW) and stored in LDS as float4
// (component k, A, amplitude, 0) for k = 0..3.  When W can be folded into A
// (lam_p ~ 1 within the error bound) the fourth K row carries (1, A) and the
// C operand is a constant zero.
#include "device_util.hpp"

namespace idisim {

typedef float v4f __attribute__((ext_vector_type(4)));

__global__ void k_reset_counter_m(int* counter) { *counter = 0; }

// R pixel rows per wave, NW waves per workgroup: tile = (R NW) x 16 pixels,
// chunks of 64 NW emitters.  FOLD: W folded into A, so the fourth K row
// carries (1, A) and C is zero; otherwise it carries (lam, W) and C = A.
template <int R, int NW, bool FOLD>
struct MfmaTile {
    static constexpr int kThreads = 64 * NW;
    static constexpr int kRows = R * NW;
    static constexpr int kChunkM = 64 * NW;

    // Accumulate emitters [e0, e1) for the tile at (tx0, ty0).  On return every
    // lane of a 16-lane row holds in tot[r][i] the full sum for pixel
    // (tx0 + R wave + r, ty0 + 4 (lane / 16) + i).
    __device__ __forceinline__ static void run(const Geom& g, const double* __restrict__ emit,
                                               const double* __restrict__ amp, int64_t e0, int64_t e1, int tx0,
                                               int ty0, float4* srec, double (&tot_re)[R][4], double (&tot_im)[R][4])
    {
        const int tid = threadIdx.x;
        const int lane = tid & 63, wave = tid >> 6;
        const int m = lane & 15, kq = lane >> 4;

        const int xe = min(tx0 + kRows, g.n0) - 1, ye = min(ty0 + 16, g.n1) - 1;
        const Frame c = pixel_frame(g, 0.5 * (tx0 + xe), 0.5 * (ty0 + ye));

        float aop[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const Frame f = pixel_frame(g, (double)(tx0 + wave * R + r), (double)(ty0 + m));
            double v;
            if (kq == 0)
                v = -g.krev * (f.u[0] - c.u[0]);
            else if (kq == 1)
                v = -g.krev * (f.u[1] - c.u[1]);
            else if (kq == 2)
                v = -g.krev * (f.u[2] - c.u[2]);
            else
                v = FOLD ? 1.0 : c.R / f.R;
            aop[r] = (float)v;
        }

        float acc_re[R][4], acc_im[R][4], cmp_re[R][4], cmp_im[R][4];
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int i = 0; i < 4; ++i) acc_re[r][i] = acc_im[r][i] = cmp_re[r][i] = cmp_im[r][i] = 0.f;

        for (int64_t base = e0; base < e1; base += kChunkM) {
            const int cnt = (e1 - base) < (int64_t)kChunkM ? (int)(e1 - base) : kChunkM;
            __syncthreads();
            {
                float4 q0 = make_float4(0.f, 0.f, 0.f, 0.f), q1 = q0, q2 = q0, q3 = q0;
                if (tid < cnt) {
                    const int64_t e = base + tid;
                    const double* a4 = emit + 4 * e;
                    const double ea = (amp != nullptr) ? amp[e] : 1.0;
                    const SplitRec rc = make_split_rec(g, c, a4[0], a4[1], a4[2], a4[3], ea);
                    // FOLD: (component, amplitude) with A in the fourth row; else (component, A, amplitude)
                    q0 = make_float4(rc.x, rc.A, rc.amp, 0.f);
                    q1 = make_float4(rc.y, rc.A, rc.amp, 0.f);
                    q2 = make_float4(rc.z, rc.A, rc.amp, 0.f);
                    q3 = make_float4(FOLD ? rc.A : rc.W, rc.A, rc.amp, 0.f);
                }
                // emitters past cnt get amplitude 0 and contribute nothing
                srec[tid] = q0;
                srec[kChunkM + tid] = q1;
                srec[2 * kChunkM + tid] = q2;
                srec[3 * kChunkM + tid] = q3;
            }
            __syncthreads();

            float pre[R][4], pim[R][4];
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int i = 0; i < 4; ++i) pre[r][i] = pim[r][i] = 0.f;

            const int nblk = (cnt + 15) >> 4;
            const float4* rowp = srec + kq * kChunkM + m;
            const v4f zero = {0.f, 0.f, 0.f, 0.f};
            // Software pipeline over pairs of blocks (ping-pong): the MFMAs of the
            // next block and the LDS reads two blocks ahead are issued before the
            // sin/cos of the current block, so the matrix-core and LDS latencies
            // overlap VALU work.  Records past cnt are zero (amplitude 0) and the
            // LDS chunk holds kChunkM/16 (even) blocks, so an even block count is safe.
            const int npair = (nblk + 1) >> 1;
            float4 qa = rowp[0], qb = rowp[16];
            v4f ta[R], tb[R];
#pragma unroll
            for (int r = 0; r < R; ++r)
                ta[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(aop[r], qa.x, FOLD ? zero : (v4f){qa.y, qa.y, qa.y, qa.y},
                                                             0, 0, 0);
            for (int p = 0; p < npair; ++p) {
#pragma unroll
                for (int r = 0; r < R; ++r)
                    tb[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(aop[r], qb.x,
                                                                 FOLD ? zero : (v4f){qb.y, qb.y, qb.y, qb.y}, 0, 0, 0);
                const float ampa = qa.z, ampb = qb.z;
                // next pair (the last iteration re-reads pair 0: in bounds, result unused)
                const int nx = (p + 1 < npair) ? 2 * p + 2 : 0;
                qa = rowp[16 * nx];
                qb = rowp[16 * (nx + 1)];
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const float sn = __builtin_amdgcn_sinf(ta[r][i]);
                        const float cs = __builtin_amdgcn_cosf(ta[r][i]);
                        pre[r][i] = fmaf(ampa, cs, pre[r][i]);
                        pim[r][i] = fmaf(ampa, sn, pim[r][i]);
                    }
#pragma unroll
                for (int r = 0; r < R; ++r)
                    ta[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(aop[r], qa.x,
                                                                 FOLD ? zero : (v4f){qa.y, qa.y, qa.y, qa.y}, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const float sn = __builtin_amdgcn_sinf(tb[r][i]);
                        const float cs = __builtin_amdgcn_cosf(tb[r][i]);
                        pre[r][i] = fmaf(ampb, cs, pre[r][i]);
                        pim[r][i] = fmaf(ampb, sn, pim[r][i]);
                    }
            }
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    kahan_add(acc_re[r][i], cmp_re[r][i], pre[r][i]);
                    kahan_add(acc_im[r][i], cmp_im[r][i], pim[r][i]);
                }
        }

        // sum the 16 emitter lanes of every 16-lane row in FP64
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                double vr = (double)acc_re[r][i] - (double)cmp_re[r][i];
                double vi = (double)acc_im[r][i] - (double)cmp_im[r][i];
#pragma unroll
                for (int off = 8; off > 0; off >>= 1) {
                    vr += __shfl_xor(vr, off, 16);
                    vi += __shfl_xor(vi, off, 16);
                }
                tot_re[r][i] = vr;
                tot_im[r][i] = vi;
            }
    }

    // Call fn(x, y, re, im) for every pixel of the tile, from the one lane of its
    // 16-lane row whose index matches the slot.
    template <typename StoreFn>
    __device__ __forceinline__ static void store(int tx0, int ty0, const double (&tot_re)[R][4],
                                                 const double (&tot_im)[R][4], StoreFn fn)
    {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int m = lane & 15, kq = lane >> 4;
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (m == (r * 4 + i) % 16) fn(tx0 + wave * R + r, ty0 + 4 * kq + i, tot_re[r][i], tot_im[r][i]);
    }
};

// Geometry of the MFMA tile used by the library: 2 rows per wave, 8 waves.
constexpr int kMR = 2, kMNW = 8;
using TileF = MfmaTile<kMR, kMNW, true>;
using TileW = MfmaTile<kMR, kMNW, false>;
static_assert(TileF::kRows == 16, "MFMA tile is 16 x 16 pixels (tile_rows/cols(kTileMfma))");

template <class T>
__global__ __launch_bounds__(T::kThreads) void k_fwd_mfma(Geom g, const double* __restrict__ emit,
                                                          const double* __restrict__ amp, int64_t n, int tiles_y,
                                                          int64_t per_split, float4* __restrict__ partial, void* out,
                                                          int kind)
{
    __shared__ float4 srec[4 * T::kChunkM];
    const int tile = blockIdx.x, split = blockIdx.y;
    const int tx0 = (tile / tiles_y) * T::kRows, ty0 = (tile % tiles_y) * 16;
    const int64_t e0 = (int64_t)split * per_split;
    const int64_t e1 = (e0 + per_split < n) ? e0 + per_split : n;
    double tre[kMR][4], tim[kMR][4];
    T::run(g, emit, amp, e0, e1, tx0, ty0, srec, tre, tim);
    const size_t npix = (size_t)g.n0 * g.n1;
    T::store(tx0, ty0, tre, tim, [&](int x, int y, double re, double im) {
        if (x < g.n0 && y < g.n1) {
            const size_t idx = (size_t)x * g.n1 + y;
            if (partial != nullptr) {
                // (hi, lo) with hi - lo = value, the encoding k_finalize_split expects
                const float hr = (float)re, hi = (float)im;
                partial[(size_t)split * npix + idx] =
                    make_float4(hr, hi, (float)((double)hr - re), (float)((double)hi - im));
            } else {
                store_split(g, x, y, re, im, out, kind, idx);
            }
        }
    });
}

template <class T>
__global__ __launch_bounds__(T::kThreads) void k_batched_mfma(Geom g, const double* __restrict__ emit,
                                                              const double* __restrict__ amp,
                                                              const WorkItem* __restrict__ items, int nitems,
                                                              int* __restrict__ counter, int tiles_y, void* out,
                                                              int kind)
{
    __shared__ float4 srec[4 * T::kChunkM];
    __shared__ int s_item;
    const size_t npix = (size_t)g.n0 * g.n1;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = s_item;
        if (it >= nitems) break;
        const WorkItem w = items[it];
        const int tx0 = (w.tile / tiles_y) * T::kRows, ty0 = (w.tile % tiles_y) * 16;
        double tre[kMR][4], tim[kMR][4];
        T::run(g, emit, amp, w.e0, w.e1, tx0, ty0, srec, tre, tim);
        char* outs = static_cast<char*>(out) +
                     (size_t)w.shot * npix * (kind == OUT_INTENSITY ? sizeof(float) : sizeof(float2));
        T::store(tx0, ty0, tre, tim, [&](int x, int y, double re, double im) {
            if (x < g.n0 && y < g.n1) store_split(g, x, y, re, im, outs, kind, (size_t)x * g.n1 + y);
        });
    }
}

hipError_t launch_fwd_mfma(const Geom& g, const double* emit, const double* amp, int64_t n, int splits,
                           float4* partial, void* out, int kind, hipStream_t st)
{
    const int tiles_x = (g.n0 + TileF::kRows - 1) / TileF::kRows, tiles_y = (g.n1 + 15) / 16;
    const int64_t ch = TileF::kChunkM;
    const int64_t per = ((n + splits - 1) / splits + ch - 1) / ch * ch;
    dim3 grid(tiles_x * tiles_y, splits);
    if (g.foldw)
        hipLaunchKernelGGL(k_fwd_mfma<TileF>, grid, dim3(TileF::kThreads), 0, st, g, emit, amp, n, tiles_y,
                           per > 0 ? per : ch, partial, out, kind);
    else
        hipLaunchKernelGGL(k_fwd_mfma<TileW>, grid, dim3(TileW::kThreads), 0, st, g, emit, amp, n, tiles_y,
                           per > 0 ? per : ch, partial, out, kind);
    return hipGetLastError();
}

hipError_t launch_batched_mfma(const Geom& g, const double* emit, const double* amp, const WorkItem* items,
                               int nitems, int* counter, int nblocks, void* out, int kind, hipStream_t st)
{
    const int tiles_y = (g.n1 + 15) / 16;
    hipLaunchKernelGGL(k_reset_counter_m, dim3(1), dim3(1), 0, st, counter);
    if (g.foldw)
        hipLaunchKernelGGL(k_batched_mfma<TileF>, dim3(nblocks), dim3(TileF::kThreads), 0, st, g, emit, amp, items,
                           nitems, counter, tiles_y, out, kind);
    else
        hipLaunchKernelGGL(k_batched_mfma<TileW>, dim3(nblocks), dim3(TileW::kThreads), 0, st, g, emit, amp, items,
                           nitems, counter, tiles_y, out, kind);
    return hipGetLastError();
}

int occupancy_batched_mfma(bool fold)
{
    int nb = 0;
    hipError_t e = fold ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batched_mfma<TileF>, TileF::kThreads, 0)
                        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batched_mfma<TileW>, TileW::kThreads, 0);
    if (e != hipSuccess) return 1;
    return nb > 0 ? nb : 1;
}

}  // namespace idisim
