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
// keys on FP32-split handles (the reference's "mixed" precision, which also takes the sine
// and cosine of the range-reduced phase in FP32) and float64 on FP64 handles.
//
// Up to 20480 emitters (float32 keys; 12288 with float64 keys): k_time_fused, one workgroup
// of 256 or 512 lanes per pixel, all in registers and LDS.
// Larger counts, per batch of pixels:
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
constexpr int kTimeK = 8;           // sorted arrivals per lane and chunk (batched path)
// One workgroup per pixel: 256 lanes with up to 40 arrivals each, or 512 lanes with up to 40
// (float32 keys) or 24 (float64 keys) arrivals each, as much as 160 KB of LDS holds.
constexpr int kFusedMax32 = 512 * 40;
constexpr int kFusedMax64 = 512 * 24;
}

namespace {
// Arrival time t and amplitude (re, im) of emitter e at pixel r (|r| = R), FP64.
// F32AMP (float32 arrival keys, the reference's mixed precision): the phase is
// reduced to revolutions in FP64 and its sine and cosine taken in FP32
// (v_sin_f32 / v_cos_f32, ~1e-7 absolute), as the reference's single path takes
// __sincosf of a float phase.
template <bool F32AMP>
__device__ inline void arrival(const Geom& g, const double r[3], double R, const double* e, double T, double c,
                               double* t, double* re, double* im)
{
    double ph, scale = 1.0;
    if (g.nf) {
        const double dx = e[0] - r[0], dy = e[1] - r[1], dz = e[2] - r[2];
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        // d - R without cancellation: d^2 - R^2 = |r_j|^2 - 2 r_j . r_p
        const double e2 = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
        const double er = e[0] * r[0] + e[1] * r[1] + e[2] * r[2];
        const double s = (e2 - 2.0 * er) / (d + R) + e[2];
        *t = s / c + T;
        ph = s * g.k + e[3];
        if (g.scale) scale = 1.0 / d;
    } else {
        const double rd = 1.0 / R;
        const double kr = g.k * rd;
        const double s = e[0] * (r[0] * kr) + e[1] * (r[1] * kr) + e[2] * (g.k * (r[2] * rd - 1.0));
        *t = -s / (c * g.k) + T;
        ph = e[3] - s;
        if (g.scale) scale = rd;
    }
    if (F32AMP) {
        double rv = ph * kInv2Pi;
        rv -= rint(rv);
        const float fr = (float)rv;
        *re = (double)__builtin_amdgcn_cosf(fr) * scale;
        *im = (double)__builtin_amdgcn_sinf(fr) * scale;
    } else {
        double sn, cs;
        sincos(ph, &sn, &cs);
        *re = cs * scale;
        *im = sn * scale;
    }
}

// The arrival time alone (the same operations as arrival(); its unused sincos is dead code).
__device__ inline double arrival_time(const Geom& g, const double r[3], double R, const double* e, double T,
                                      double c)
{
    double t, re, im;
    arrival<false>(g, r, R, e, T, c, &t, &re, &im);
    return t;
}

// The amplitude alone (the same operations as arrival()).
template <bool F32AMP>
__device__ inline void arrival_amp(const Geom& g, const double r[3], double R, const double* e, double* re,
                                   double* im)
{
    double t;
    arrival<F32AMP>(g, r, R, e, 0.0, 1.0, &t, re, im);
}
}  // namespace

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
    const double R = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    double t, re, im;
    arrival<sizeof(Tt) == 4>(g, r, R, emit + 4 * j, times[j], c, &t, &re, &im);
    const size_t o = (size_t)pl * n + j;
    keys[o] = (Tt)t;
    idx[o] = (int)j;
    amp[o] = make_double2(re, im);
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

// Small emitter counts (n <= 256 ITEMS): one workgroup per pixel does everything in LDS.
// The arrival keys are formed in a plain loop into LDS, moved into registers (blocked: lane l
// holds emitters l ITEMS .. l ITEMS + ITEMS - 1) for rocPRIM's block radix sort (stable) and
// written back sorted, with the emitter indices, to a padded LDS array (odd stride: no bank
// conflicts).  The amplitudes are recomputed in sorted order: once for the lanes' aggregate
// states (scanned across lanes as above) and once more to step the field through the lane's
// arrivals from its incoming state and sum the integral terms.  The key buffer, the sort's
// storage and the sorted arrays share one LDS union; nothing goes through global memory.
template <typename Tt, int ITEMS, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_time_fused(Geom g, const double* __restrict__ emit,
                                                    const double* __restrict__ times, int64_t n, double c,
                                                    double tau, double* __restrict__ out)
{
    using Sort = rocprim::block_radix_sort<Tt, BLOCK, ITEMS, int>;
    constexpr int S = ITEMS + 1;  // padded stride of the sorted arrays
    static_assert(BLOCK * ITEMS <= 65536, "sorted indices are stored as 16 bits");
    union Lds {
        typename Sort::storage_type sort;
        Tt kbuf[BLOCK * ITEMS];
        struct {
            Tt key[BLOCK * S];
            unsigned short idx[BLOCK * S];
        } sorted;
    };
    __shared__ Lds u;
    __shared__ double sre[BLOCK], sim[BLOCK], st[BLOCK];
    __shared__ int shas[BLOCK];
    __shared__ double ssum[BLOCK];
    const int tid = threadIdx.x;
    const int64_t p = blockIdx.x;
    const double inv_tau = 1.0 / tau, two_over_tau = 2.0 / tau;
    double r[3];
    pixel_pos(g, (double)(p / g.n1), (double)(p % g.n1), r);
    const double R = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);

    // arrival keys (emitters in order, coalesced reads)
#pragma unroll 1
    for (int j = tid; j < BLOCK * ITEMS; j += BLOCK) {
        Tt key = (Tt)INFINITY;
        if (j < n) key = (Tt)arrival_time(g, r, R, emit + 4 * (int64_t)j, times[j], c);
        u.kbuf[j] = key;
    }
    __syncthreads();
    Tt keys[ITEMS];
    int js[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        keys[k] = u.kbuf[tid * ITEMS + k];
        js[k] = tid * ITEMS + k;
    }
    __syncthreads();  // kbuf read by every lane before the sort reuses the storage
    Sort().sort(keys, js, u.sort);
    __syncthreads();  // sort storage consumed before it becomes the sorted arrays
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        u.sorted.key[tid * S + k] = keys[k];
        u.sorted.idx[tid * S + k] = (unsigned short)js[k];
    }
    __syncthreads();  // every lane reads its neighbours' first and last keys
    const Tt* skey = u.sorted.key + tid * S;
    const unsigned short* sidx = u.sorted.idx + tid * S;
    const int64_t i0 = (int64_t)tid * ITEMS;
    const int cnt = (int)((n - i0) < ITEMS ? ((n - i0) > 0 ? n - i0 : 0) : ITEMS);

    // One pass over this lane's sorted arrivals.  With x_in the field just after the
    // previous arrival (time tprev, known only after the scan) and P_k the lane's own
    // field (x_in = 0), x_k = P_k + D_k x_in with D_k = e^{-(t_k - tprev)/tau}, so
    //   sum_k w_k |x_k|^2 = A + 2 Re(conj(x_in) B) + |x_in|^2 C,
    //   A = sum w |P|^2,  B = sum w D P,  C = sum w D^2,
    // and the lane's aggregate (P_last, t_last) feeds the scan.  The previous lane is
    // full whenever this one has arrivals, so its last key is tprev.
    const double tprev = (cnt > 0) ? (double)((i0 > 0) ? u.sorted.key[(tid - 1) * S + ITEMS - 1] : skey[0]) : 0.0;
    double pr = 0.0, pi = 0.0, D = 1.0, tl = tprev;
    double A = 0.0, Br = 0.0, Bi = 0.0, C = 0.0;
#pragma unroll 1
    for (int k = 0; k < cnt; ++k) {
        const int j = sidx[k];
        double re, im;
        arrival_amp<sizeof(Tt) == 4>(g, r, R, emit + 4 * (int64_t)j, &re, &im);
        const double tk = (double)skey[k];
        const double dk = exp(-(tk - tl) * inv_tau);
        pr = re + dk * pr;
        pi = im + dk * pi;
        D *= dk;
        double w = 1.0;
        if (i0 + k + 1 < n) {
            const double tn = (k + 1 < ITEMS) ? (double)skey[k + 1] : (double)u.sorted.key[(tid + 1) * S];
            w = -expm1((tk - tn) * two_over_tau);
        }
        A += w * (pr * pr + pi * pi);
        const double wD = w * D;
        Br += wD * pr;
        Bi += wD * pi;
        C += wD * D;
        tl = tk;
    }
    const DecayState mine = {pr, pi, tl, cnt > 0 ? 1 : 0};
    sre[tid] = mine.re;
    sim[tid] = mine.im;
    st[tid] = mine.t;
    shas[tid] = mine.has;
    __syncthreads();
    DecayState inc = mine;
    for (int off = 1; off < BLOCK; off <<= 1) {
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
    double xr = 0.0, xi = 0.0;
    if (tid > 0 && shas[tid - 1]) {
        xr = sre[tid - 1];
        xi = sim[tid - 1];
    }
    const double acc = A + 2.0 * (xr * Br + xi * Bi) + (xr * xr + xi * xi) * C;
    ssum[tid] = acc;
    __syncthreads();
    for (int s2 = BLOCK / 2; s2 > 0; s2 >>= 1) {
        if (tid < s2) ssum[tid] += ssum[tid + s2];
        __syncthreads();
    }
    if (tid == 0) out[p] = ssum[0] * (0.5 * tau);
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

template <typename Tt>
static hipError_t run_fused(const Geom& g, const double* emit, const double* times, int64_t n, double c, double tau,
                            double* out, int64_t npix, hipStream_t st)
{
    const dim3 grid((unsigned)npix);
#define IDISIM_TIME_FUSED(I, B) \
    hipLaunchKernelGGL((k_time_fused<Tt, I, B>), grid, dim3(B), 0, st, g, emit, times, n, c, tau, out)
    if (n <= 256 * 4)
        IDISIM_TIME_FUSED(4, 256);
    else if (n <= 256 * 8)
        IDISIM_TIME_FUSED(8, 256);
    else if (n <= 256 * 16)
        IDISIM_TIME_FUSED(16, 256);
    else if (n <= 256 * 24)
        IDISIM_TIME_FUSED(24, 256);
    else if (n <= 256 * 32)
        IDISIM_TIME_FUSED(32, 256);
    else if (n <= 256 * 40)
        IDISIM_TIME_FUSED(40, 256);
    else if (n <= 512 * 24)
        IDISIM_TIME_FUSED(24, 512);
    else {
        if constexpr (sizeof(Tt) == 4) {  // float64 keys: 512 x 32 no longer fits the LDS
            if (n <= 512 * 32)
                IDISIM_TIME_FUSED(32, 512);
            else if (n <= 512 * 40)
                IDISIM_TIME_FUSED(40, 512);
            else
                return hipErrorInvalidValue;  // past the fused capacity (time_fused_ok)
        } else {
            return hipErrorInvalidValue;
        }
    }
#undef IDISIM_TIME_FUSED
    return hipGetLastError();
}

__global__ void k_time_zero(double* __restrict__ out, int64_t npix)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < npix) out[i] = 0.0;
}

bool time_fused_ok(int64_t n, bool fp64_keys, int path)
{
    return path != 2 && n <= (fp64_keys ? kFusedMax64 : kFusedMax32);
}

hipError_t launch_time(const Geom& g, const TimeLayout& L, const double* emit, const double* times, int64_t n,
                       double c, double tau, int path, void* scratch, double* out, hipStream_t st)
{
    if (n == 0) {
        const int64_t npix = (int64_t)g.n0 * g.n1;
        hipLaunchKernelGGL(k_time_zero, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, st, out, npix);
        return hipGetLastError();
    }
    if (time_fused_ok(n, L.fp64, path)) {
        const int64_t npix = (int64_t)g.n0 * g.n1;
        if (L.fp64) return run_fused<double>(g, emit, times, n, c, tau, out, npix, st);
        return run_fused<float>(g, emit, times, n, c, tau, out, npix, st);
    }
    return L.fp64 ? run_time<double>(g, L, emit, times, n, c, tau, scratch, out, st)
                  : run_time<float>(g, L, emit, times, n, c, tau, scratch, out, st);
}

}  // namespace idisim
