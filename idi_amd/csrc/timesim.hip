// Time-dependent simulation with decaying emitters (reference: idi/simulation/time.py:7-98 on
// the CPU, cutime.py:9-138 and simtime.cu:43-242 on the GPU).
//
// Emitter j, excited at time T_j, radiates a_j e^{-(t - t_j)/tau} from its arrival time t_j
// at pixel p on.  The pixel records
//   I_p = integral |E_p(t)|^2 dt,   E_p(t) = sum_{t_j <= t} a_j e^{-(t - t_j)/tau}.
// With the arrivals sorted, x_i = a_i + e^{-(t_i - t_{i-1})/tau} x_{i-1} is the field just
// after arrival i and
//   I_p = tau/2 [ sum_{i<n} -|x_i|^2 expm1(-2 (t_{i+1} - t_i)/tau) + |x_n|^2 ].
// Arrival time and amplitude per (pixel, emitter) (simtime.cu:43-112):
//   near field: s = |r_j - r_p| - |r_p| + z_j, t = s/c + T_j, a = e^{i(k s + phi_j)} (/ |r_j - r_p|)
//   far field:  s = q_p . r_j with q_p = k (r_p/|r_p| - z_hat), t = -s/(c k) + T_j,
//               a = e^{i(phi_j - s)} (/ |r_p|)
// in FP64 (d - |r_p| as (|r_j|^2 - 2 r_j . r_p)/(d + |r_p|), free of the cancellation of
// two ~detector-distance terms that the reference's formula has); arrival times are float32
// keys on FP32-split handles (the reference's "mixed" precision) and float64 on FP64 handles.
//
// Work per batch of pixels:
//   k_time_prepare    one thread per (pixel, emitter): key t, index j, amplitude a (FP64);
//   segmented radix sort of (t, j) per pixel (rocPRIM, stable);
//   k_time_integrate  one workgroup per pixel: 8 sorted arrivals per lane, decayed partial
//                     sums per lane, a Hillis-Steele scan of the lanes' states under
//                     (t_A, X_A) o (t_B, X_B) = (t_B, X_B + e^{-(t_B - t_A)/tau} X_A), the
//                     carry between chunks of 2048, and the integral terms summed in FP64.
#include <cstring>  // host memset, used by rocPRIM headers

#include <rocprim/rocprim.hpp>

#include "kernels.hpp"

namespace idisim {

namespace {
constexpr int kTimeK = 8;  // sorted arrivals per lane and chunk
}

template <typename Tt>
__global__ __launch_bounds__(256) void k_time_prepare(Geom g, const double* __restrict__ emit,
                                                      const double* __restrict__ times, int64_t n, int64_t p0,
                                                      double c, Tt* __restrict__ keys, int* __restrict__ idx,
                                                      double2* __restrict__ amp)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int64_t pl = blockIdx.y;
    const int64_t p = p0 + pl;
    double r[3];
    pixel_pos(g, (double)(p / g.n1), (double)(p % g.n1), r);
    const double* e = emit + 4 * j;
    double t, ph, scale = 1.0;
    if (g.nf) {
        const double dx = e[0] - r[0], dy = e[1] - r[1], dz = e[2] - r[2];
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        const double R = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        // d - R without cancellation: d^2 - R^2 = |r_j|^2 - 2 r_j . r_p
        const double e2 = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
        const double er = e[0] * r[0] + e[1] * r[1] + e[2] * r[2];
        const double s = (e2 - 2.0 * er) / (d + R) + e[2];
        t = s / c + times[j];
        ph = s * g.k + e[3];
        if (g.scale) scale = 1.0 / d;
    } else {
        const double rd = 1.0 / sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        const double kr = g.k * rd;
        const double s = e[0] * (r[0] * kr) + e[1] * (r[1] * kr) + e[2] * (g.k * (r[2] * rd - 1.0));
        t = -s / (c * g.k) + times[j];
        ph = e[3] - s;
        if (g.scale) scale = rd;
    }
    double sn, cs;
    sincos(ph, &sn, &cs);
    const size_t o = (size_t)pl * n + j;
    keys[o] = (Tt)t;
    idx[o] = (int)j;
    amp[o] = make_double2(cs * scale, sn * scale);
}

__global__ void k_time_offsets(int* __restrict__ offs, int np, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= np) offs[i] = i * n;
}

namespace {
struct DecayState {
    double re, im, t;
    int has;
};

// (t_A, X_A) o (t_B, X_B) = (t_B, X_B + e^{-(t_B - t_A)/tau} X_A); an empty state is the identity
__device__ inline DecayState combine(const DecayState& a, const DecayState& b, double inv_tau)
{
    if (!a.has) return b;
    if (!b.has) return a;
    const double d = exp(-(b.t - a.t) * inv_tau);
    DecayState r;
    r.re = b.re + d * a.re;
    r.im = b.im + d * a.im;
    r.t = b.t;
    r.has = 1;
    return r;
}
}  // namespace

template <typename Tt>
__global__ __launch_bounds__(256) void k_time_integrate(const Tt* __restrict__ keys, const int* __restrict__ idx,
                                                        const double2* __restrict__ amp, int64_t n, double tau,
                                                        int64_t p0, double* __restrict__ out)
{
    __shared__ double sre[256], sim[256], st[256];
    __shared__ int shas[256];
    __shared__ double ssum[256];
    const int tid = threadIdx.x;
    const int64_t pl = blockIdx.x;
    const Tt* t = keys + pl * n;
    const int* id = idx + pl * n;
    const double2* a = amp + pl * n;
    const double inv_tau = 1.0 / tau, two_over_tau = 2.0 / tau;

    DecayState carry = {0.0, 0.0, 0.0, 0};
    double acc = 0.0;
    for (int64_t c0 = 0; c0 < n; c0 += 256 * kTimeK) {
        const int64_t b = c0 + (int64_t)tid * kTimeK;
        // this lane's arrivals, decayed among themselves (as if nothing came before)
        double lre[kTimeK], lim[kTimeK], lt[kTimeK];
        DecayState mine = {0.0, 0.0, 0.0, 0};
#pragma unroll
        for (int k = 0; k < kTimeK; ++k) {
            const int64_t i = b + k;
            if (i < n) {
                const double ti = (double)t[i];
                const double2 ai = a[id[i]];
                DecayState s = {ai.x, ai.y, ti, 1};
                mine = combine(mine, s, inv_tau);
            }
            lre[k] = mine.re;
            lim[k] = mine.im;
            lt[k] = mine.t;
        }
        // inclusive scan of the lanes' states (Hillis-Steele in LDS)
        sre[tid] = mine.re;
        sim[tid] = mine.im;
        st[tid] = mine.t;
        shas[tid] = mine.has;
        __syncthreads();
        DecayState inc = mine;
        for (int off = 1; off < 256; off <<= 1) {
            DecayState left = {0.0, 0.0, 0.0, 0};
            if (tid >= off) left = {sre[tid - off], sim[tid - off], st[tid - off], shas[tid - off]};
            __syncthreads();
            inc = combine(left, inc, inv_tau);
            sre[tid] = inc.re;
            sim[tid] = inc.im;
            st[tid] = inc.t;
            shas[tid] = inc.has;
            __syncthreads();
        }
        DecayState before = {0.0, 0.0, 0.0, 0};
        if (tid > 0) before = {sre[tid - 1], sim[tid - 1], st[tid - 1], shas[tid - 1]};
        const DecayState in = combine(carry, before, inv_tau);
        // fields after each of this lane's arrivals and the integral terms
#pragma unroll
        for (int k = 0; k < kTimeK; ++k) {
            const int64_t i = b + k;
            if (i >= n) break;
            double xr = lre[k], xi = lim[k];
            if (in.has) {
                const double d = exp(-(lt[k] - in.t) * inv_tau);
                xr += d * in.re;
                xi += d * in.im;
            }
            const double m2 = xr * xr + xi * xi;
            if (i + 1 < n) {
                const double tn = (double)t[i + 1];
                acc += -m2 * expm1((lt[k] - tn) * two_over_tau);
            } else {
                acc += m2;
            }
        }
        // carry into the next chunk: the state after the chunk's last arrival
        const DecayState total = combine(carry, {sre[255], sim[255], st[255], shas[255]}, inv_tau);
        __syncthreads();
        carry = total;
    }
    ssum[tid] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) ssum[tid] += ssum[tid + s];
        __syncthreads();
    }
    if (tid == 0) out[p0 + pl] = ssum[0] * (0.5 * tau);
}

template <typename Tt>
static size_t time_sort_bytes(int np, int64_t n)
{
    size_t bytes = 0;
    const size_t total = (size_t)np * (size_t)n;
    (void)rocprim::segmented_radix_sort_pairs((void*)nullptr, bytes, (const Tt*)nullptr, (Tt*)nullptr,
                                              (const int*)nullptr, (int*)nullptr, total, (unsigned)np,
                                              (const int*)nullptr, (const int*)nullptr, 0, 8 * sizeof(Tt));
    return bytes;
}

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

TimeLayout time_layout(const Geom& g, int64_t n, bool fp64_keys, size_t budget)
{
    TimeLayout L;
    const int64_t npix = (int64_t)g.n0 * g.n1;
    const size_t kb = fp64_keys ? 8 : 4;
    const size_t per_pixel = (size_t)n * (2 * kb + 2 * sizeof(int) + sizeof(double2)) + 1024;
    int64_t np = (int64_t)(budget / (per_pixel > 0 ? per_pixel : 1));
    if (np > npix) np = npix;
    if (np < 1) np = 1;
    if (np > 65535) np = 65535;                               // grid.y of the prepare kernel
    while (np > 1 && np * n >= ((int64_t)1 << 31)) np /= 2;  // 32-bit segment offsets
    L.np = (int)np;
    L.fp64 = fp64_keys;
    const size_t total = (size_t)np * (size_t)n;
    L.keys_bytes = align256(total * kb);
    L.idx_bytes = align256(total * sizeof(int));
    L.amp_bytes = align256(total * sizeof(double2));
    L.off_bytes = align256((size_t)(np + 1) * sizeof(int));
    L.sort_bytes = align256(fp64_keys ? time_sort_bytes<double>(L.np, n) : time_sort_bytes<float>(L.np, n));
    L.bytes = 2 * L.keys_bytes + 2 * L.idx_bytes + L.amp_bytes + L.off_bytes + L.sort_bytes;
    return L;
}

template <typename Tt>
static hipError_t run_time(const Geom& g, const TimeLayout& L, const double* emit, const double* times, int64_t n,
                           double c, double tau, void* scratch, double* out, hipStream_t st)
{
    char* base = static_cast<char*>(scratch);
    Tt* k0 = reinterpret_cast<Tt*>(base);
    Tt* k1 = reinterpret_cast<Tt*>(base + L.keys_bytes);
    int* i0 = reinterpret_cast<int*>(base + 2 * L.keys_bytes);
    int* i1 = reinterpret_cast<int*>(base + 2 * L.keys_bytes + L.idx_bytes);
    double2* amp = reinterpret_cast<double2*>(base + 2 * L.keys_bytes + 2 * L.idx_bytes);
    int* offs = reinterpret_cast<int*>(base + 2 * L.keys_bytes + 2 * L.idx_bytes + L.amp_bytes);
    void* sort_tmp = base + 2 * L.keys_bytes + 2 * L.idx_bytes + L.amp_bytes + L.off_bytes;
    const int64_t npix = (int64_t)g.n0 * g.n1;
    for (int64_t p0 = 0; p0 < npix; p0 += L.np) {
        const int np = (int)((npix - p0) < L.np ? (npix - p0) : L.np);
        hipLaunchKernelGGL(k_time_offsets, dim3((unsigned)((np + 256) / 256)), dim3(256), 0, st, offs, np, (int)n);
        hipLaunchKernelGGL((k_time_prepare<Tt>), dim3((unsigned)((n + 255) / 256), (unsigned)np), dim3(256), 0, st, g,
                           emit, times, n, p0, c, k0, i0, amp);
        size_t bytes = L.sort_bytes;
        hipError_t e = rocprim::segmented_radix_sort_pairs(sort_tmp, bytes, k0, k1, i0, i1, (size_t)np * (size_t)n,
                                                           (unsigned)np, offs, offs + 1, 0, 8 * (int)sizeof(Tt), st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_time_integrate<Tt>), dim3((unsigned)np), dim3(256), 0, st, k1, i1, amp, n, tau, p0, out);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

__global__ void k_time_zero(double* __restrict__ out, int64_t npix)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < npix) out[i] = 0.0;
}

hipError_t launch_time(const Geom& g, const TimeLayout& L, const double* emit, const double* times, int64_t n,
                       double c, double tau, void* scratch, double* out, hipStream_t st)
{
    if (n == 0) {
        const int64_t npix = (int64_t)g.n0 * g.n1;
        hipLaunchKernelGGL(k_time_zero, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, st, out, npix);
        return hipGetLastError();
    }
    return L.fp64 ? run_time<double>(g, L, emit, times, n, c, tau, scratch, out, st)
                  : run_time<float>(g, L, emit, times, n, c, tau, scratch, out, st);
}

}  // namespace idisim
