// Detector charge sharing on the GPU (reference: chargesharing / __putcharges,
// idi/simulation/common.py:26-67).
//
// Each photon count of the input image is a hit at a uniformly random position
// mu in [-1/2, 1/2)^2 inside its pixel.  Its charge is a Gaussian of width s,
// integrated over the (2w+1)^2 pixels around it (w = ceil(2 s)); charge that
// leaves the detector is lost.  For the pixel offset d along one axis the
// integral is E(mu, d) / 2 with
//   E(mu, d) = erf(c + a) + erf(c - a),  a = (mu - d) / (sqrt(2) s),  c = 1 / (2 sqrt(2) s),
// and the deposited value is (1/4 E(mu_x, .)) E(mu_y, .).  The reference adds
// k[x, y] into ret[hit_x + y - w, hit_y + x - w] with k[x, .] built from mu_x,
// so mu_x goes with the offset along the second image axis and mu_y with the
// first; that assignment is kept.
//
// The reference scatters every hit into its neighbourhood in hit order (input
// pixels in row-major order, then hit number).  Here one thread per OUTPUT
// pixel gathers the hits of its (2w+1)^2 input neighbours in exactly that
// order, which gives the same FP64 sum without atomics and deterministically.
//
// Random numbers (Philox4x32-10, key = seed): hit h of input pixel p uses the
// block (p_lo, p_hi, h, 1) for (mu_x, mu_y) and, with gain noise ns > 0, the
// block (p_lo, p_hi, h, 2) for a Box-Muller normal z; the hit's charge is then
// scaled by 1 + z ns.  Without charge sharing (s <= 0) pixel p is multiplied by
// 1 + ns z with z from (p_lo, p_hi, 0, 3).
#include "kernels.hpp"
#include "philox.hpp"

namespace idisim {

namespace {
constexpr uint32_t kTagHit = 1u, kTagHitGain = 2u, kTagPixelGain = 3u;

__device__ inline U4 draw(uint64_t p, uint32_t h, uint32_t tag, uint32_t k0, uint32_t k1)
{
    const U4 c = {{(uint32_t)p, (uint32_t)(p >> 32), h, tag}};
    return philox4x32_10(c, k0, k1);
}

// Box-Muller: sqrt(-2 ln u1) cos(2 pi u2), u1 in (0, 1], u2 in [0, 1)
__device__ inline double normal(const U4& r)
{
    const double u1 = 1.0 - u53(r.v[0], r.v[1]);
    const double u2 = u53(r.v[2], r.v[3]);
    return sqrt(-2.0 * log(u1)) * cos(kTwoPi * u2);
}

// Integer photon count of a pixel: truncation toward zero (numpy astype(int)),
// non-positive and NaN counts give no hits, at most 2^32 - 1 hits.
__device__ inline uint32_t hits_of(double c)
{
    if (!(c >= 1.0)) return 0u;
    return c >= 4294967295.0 ? 4294967295u : (uint32_t)c;
}
}  // namespace

__global__ __launch_bounds__(256) void k_chargesharing(const double* __restrict__ counts, int n0, int n1, int w,
                                                       double s, double ns, uint32_t k0, uint32_t k1,
                                                       double* __restrict__ out)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n0 * n1) return;
    const int ox = (int)(idx / n1), oy = (int)(idx - (int64_t)ox * n1);
    const double rs = sqrt(2.0) * s;
    const double c = 1.0 / (2.0 * sqrt(2.0) * s);
    const int x0 = max(ox - w, 0), x1 = min(ox + w, n0 - 1);
    const int y0 = max(oy - w, 0), y1 = min(oy + w, n1 - 1);
    double acc = 0.0;
    for (int ix = x0; ix <= x1; ++ix) {
        const double db = (double)(ox - ix);  // offset along the first axis: goes with mu_y
        for (int iy = y0; iy <= y1; ++iy) {
            const double da = (double)(oy - iy);  // offset along the second axis: goes with mu_x
            const uint64_t p = (uint64_t)ix * (uint64_t)n1 + (uint64_t)iy;
            const uint32_t nh = hits_of(counts[p]);
            for (uint32_t h = 0; h < nh; ++h) {
                const U4 r = draw(p, h, kTagHit, k0, k1);
                const double mux = u53(r.v[0], r.v[1]) - 0.5;
                const double muy = u53(r.v[2], r.v[3]) - 0.5;
                const double a = (mux - da) / rs;
                const double b = (muy - db) / rs;
                const double k = (0.25 * (erf(c + a) + erf(c - a))) * (erf(c + b) + erf(c - b));
                if (ns > 0.0) {
                    const double scale = 1.0 + normal(draw(p, h, kTagHitGain, k0, k1)) * ns;
                    acc += scale * k;
                } else {
                    acc += k;
                }
            }
        }
    }
    out[idx] = acc;
}

// Tiled form of k_chargesharing: a workgroup of 256 lanes owns 16 x 16 output
// pixels.  The hits of the (16 + 2w)^2 input pixels around the tile are numbered in
// the reference's order (input pixels row-major, then hit number) through a scan of
// their counts.  They are taken in chunks of `cap`: every hit of a chunk is drawn
// once and its record stored in LDS, namely its 2w+1 factors 0.25 (erf(c + a) +
// erf(c - a)) along the second axis, its 2w+1 factors erf(c + b) + erf(c - b) along
// the first axis and its gain scale (the reference's own per-hit work,
// common.py:53-67).  Each lane then adds, for its input neighbours in order, the
// chunk's hits with one product per deposit.  Chunks follow the global hit order, so
// each output pixel still sums its deposits in the reference's order, with the same
// FP64 values as k_chargesharing, without redrawing a hit for every deposit.
// Dynamic LDS: (16 + 2w)^2 + 1 ints (hit offsets), then cap records of 4w + 3 doubles.
constexpr int kCsTile = 16;

__global__ __launch_bounds__(256) void k_chargesharing_tile(const double* __restrict__ counts, int n0, int n1, int w,
                                                            double s, double ns, uint32_t k0, uint32_t k1, int cap,
                                                            double* __restrict__ out)
{
    extern __shared__ double lds_cs[];
    __shared__ int part[256];
    const int hw = kCsTile + 2 * w;      // halo width
    const int nh = hw * hw;              // input pixels of the halo
    const int rl = 4 * w + 3;            // doubles per hit record
    int* off = reinterpret_cast<int*>(lds_cs);
    double* rec = lds_cs + (nh + 2) / 2;  // records after the offsets (8-byte aligned)
    const int tid = threadIdx.x;
    const int tx0 = blockIdx.y * kCsTile, ty0 = blockIdx.x * kCsTile;  // tile origin (first, second axis)
    const int hx0 = tx0 - w, hy0 = ty0 - w;

    // exclusive scan of the halo's hit counts: a contiguous segment per lane, then the lanes
    const int seg = (nh + 255) / 256;
    const int q0 = tid * seg, q1 = min(q0 + seg, nh);
    int sum = 0;
    for (int q = q0; q < q1; ++q) {
        const int ix = hx0 + q / hw, iy = hy0 + q % hw;
        int c = 0;
        if (ix >= 0 && ix < n0 && iy >= 0 && iy < n1) c = (int)hits_of(counts[(size_t)ix * n1 + iy]);
        off[q + 1] = c;
        sum += c;
    }
    part[tid] = sum;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        const int v = (tid >= d) ? part[tid - d] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = (tid > 0) ? part[tid - 1] : 0;
    if (tid == 0) off[0] = 0;
    for (int q = q0; q < q1; ++q) {
        run += off[q + 1];
        off[q + 1] = run;
    }
    __syncthreads();
    const int total = off[nh];

    const double rs = sqrt(2.0) * s;
    const double c = 1.0 / (2.0 * sqrt(2.0) * s);
    const int ox = tx0 + tid / kCsTile, oy = ty0 + tid % kCsTile;
    const bool on = ox < n0 && oy < n1;
    const int x0 = max(ox - w, 0), x1 = min(ox + w, n0 - 1);
    const int y0 = max(oy - w, 0), y1 = min(oy + w, n1 - 1);
    double acc = 0.0;
    for (int c0 = 0; c0 < total; c0 += cap) {
        const int c1 = min(total, c0 + cap);
        __syncthreads();  // the previous chunk's records have been read
        for (int j = c0 + tid; j < c1; j += 256) {
            // input pixel of hit j: the last q with off[q] <= j
            int lo = 0, hi = nh;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (off[mid] <= j) lo = mid;
                else hi = mid;
            }
            const int ix = hx0 + lo / hw, iy = hy0 + lo % hw;
            const uint32_t h = (uint32_t)(j - off[lo]);
            const uint64_t p = (uint64_t)ix * (uint64_t)n1 + (uint64_t)iy;
            double* r = rec + (size_t)(j - c0) * rl;
            const U4 d = draw(p, h, kTagHit, k0, k1);
            const double mux = u53(d.v[0], d.v[1]) - 0.5;
            const double muy = u53(d.v[2], d.v[3]) - 0.5;
            for (int t = -w; t <= w; ++t) {
                const double a = (mux - (double)t) / rs;
                const double b = (muy - (double)t) / rs;
                r[t + w] = 0.25 * (erf(c + a) + erf(c - a));
                r[2 * w + 1 + t + w] = erf(c + b) + erf(c - b);
            }
            r[rl - 1] = (ns > 0.0) ? 1.0 + normal(draw(p, h, kTagHitGain, k0, k1)) * ns : 1.0;
        }
        __syncthreads();
        if (on) {
            for (int ix = x0; ix <= x1; ++ix) {
                const int fy = (ox - ix) + w;  // offset along the first axis: goes with mu_y
                for (int iy = y0; iy <= y1; ++iy) {
                    const int fx = (oy - iy) + w;  // offset along the second axis: goes with mu_x
                    const int q = (ix - hx0) * hw + (iy - hy0);
                    const int h0 = max(off[q], c0), h1 = min(off[q + 1], c1);
                    for (int h = h0; h < h1; ++h) {
                        const double* r = rec + (size_t)(h - c0) * rl;
                        const double k = r[fx] * r[2 * w + 1 + fy];
                        acc += r[rl - 1] * k;
                    }
                }
            }
        }
    }
    if (on) out[(size_t)ox * n1 + oy] = acc;
}

__global__ __launch_bounds__(256) void k_pixel_gain(const double* __restrict__ counts, int64_t npix, double ns,
                                                    uint32_t k0, uint32_t k1, double* __restrict__ out)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    out[p] = (ns > 0.0) ? counts[p] * (1.0 + ns * normal(draw((uint64_t)p, 0u, kTagPixelGain, k0, k1))) : counts[p];
}

hipError_t launch_chargesharing(const double* counts, int n0, int n1, double s, double ns, uint64_t seed,
                                double* out, hipStream_t st)
{
    const int64_t npix = (int64_t)n0 * n1;
    if (npix <= 0) return hipSuccess;
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const dim3 grid((unsigned)((npix + 255) / 256));
    if (s > 0.0) {
        const int w = (int)std::ceil(2.0 * s);
        if (w <= 8) {
            // tiled: chunks of hit records in LDS (48 KB of them)
            const int hw = kCsTile + 2 * w;
            const size_t off_bytes = (size_t)((hw * hw + 2) / 2) * sizeof(double);
            const int rl = 4 * w + 3;
            const int cap = (int)((48 * 1024) / ((size_t)rl * sizeof(double)));
            const dim3 tgrid((unsigned)((n1 + kCsTile - 1) / kCsTile), (unsigned)((n0 + kCsTile - 1) / kCsTile));
            hipLaunchKernelGGL(k_chargesharing_tile, tgrid, dim3(256), off_bytes + (size_t)cap * rl * sizeof(double),
                               st, counts, n0, n1, w, s, ns, k0, k1, cap, out);
        } else {
            hipLaunchKernelGGL(k_chargesharing, grid, dim3(256), 0, st, counts, n0, n1, w, s, ns, k0, k1, out);
        }
    } else {
        hipLaunchKernelGGL(k_pixel_gain, grid, dim3(256), 0, st, counts, npix, ns, k0, k1, out);
    }
    return hipGetLastError();
}

}  // namespace idisim
