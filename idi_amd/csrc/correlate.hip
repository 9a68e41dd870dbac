// Correlators of photon images (reference: idi/reconstruction/cusimple.py:44-121
// and cucor.py:9-133).
//
// k_corr_accumulate: one streaming step of correlator2d.add (cusimple.py:80-107):
//   c = corr / maskcorr where maskcorr > 0.5, else 0;
//   count == 1: mean = c, m2 = 0;  else: d = c - mean; mean += d / count; m2 += d (c - mean)
// (Welford's running mean and sum of squared deviations), fused into one pass, in
// float32 with the reference's operation order and no contraction.
//
// k_qcorr: q-corrected autocorrelation (cucor.py:54-91).  For a detector z pixels
// away, pixel (r, c) has zd = z / |(r - xc, c - yc, z)|; the pair (ref, target) with
// row offset 0 <= dr < qmax and column offset |dc| < qmax is binned at
//   dqx = round((tr - xc) zd_t - (rr - xc) zd_r),  dqy = round((tc - yc) zd_t - (rc - yc) zd_r),
//   dqz = round(z (zd_r - zd_t))
// (float32 arithmetic, no contraction, round half to even) and adds img_r img_t.
// One workgroup per (reference row, row offset); the two rows are staged in LDS.
// Sums are FP64 atomics (the reference accumulates float32); bins outside the
// output are skipped (the reference does not check them).
// k_qcorr_assemble (cucor.py:93-96): out[i] and out[mirror(i)], mirror(i) = (-i) mod
// shape, both become their sum for every bin with dqx index < qmax (a bin that is
// its own mirror is doubled).  In the reference two threads update each pair of
// the dqx index 0 column at once; here that pair also gets the plain sum.
#include <algorithm>

#include "kernels.hpp"

namespace idisim {

__global__ __launch_bounds__(256) void k_corr_accumulate(const float* __restrict__ corr,
                                                         const float* __restrict__ maskcorr, int64_t n,
                                                         int64_t count, float* __restrict__ mean,
                                                         float* __restrict__ m2)
{
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float mc = maskcorr[i];
    const float c = (mc > 0.5f) ? corr[i] / mc : 0.f;
    if (count <= 1) {
        mean[i] = c;
        m2[i] = 0.f;
    } else {
        const float m = mean[i];
        const float d = c - m;
        const float mn = m + d / (float)count;
        mean[i] = mn;
        m2[i] += d * (c - mn);
    }
}

__global__ __launch_bounds__(256) void k_fill_f64(double* __restrict__ p, int64_t n, double v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_qcorr(const float* __restrict__ zd, const float* __restrict__ img, int nr,
                                               int nc, float xc, float yc, float cz, int qmax, int maxdqz,
                                               double* __restrict__ out)
{
#pragma clang fp contract(off)
    extern __shared__ float lds[];  // 4 * nc floats when staged
    const int refr = blockIdx.x, offr = blockIdx.y;
    const int tr = refr + offr;
    if (tr >= nr) return;  // whole workgroup: uniform
    const bool staged = nc <= 1024;
    const float *rz, *rv, *tz, *tv;
    if (staged) {
        float* srz = lds;
        float* srv = lds + nc;
        float* stz = lds + 2 * nc;
        float* stv = lds + 3 * nc;
        for (int i = threadIdx.x; i < nc; i += blockDim.x) {
            srz[i] = zd[(size_t)refr * nc + i];
            srv[i] = img[(size_t)refr * nc + i];
            stz[i] = zd[(size_t)tr * nc + i];
            stv[i] = img[(size_t)tr * nc + i];
        }
        __syncthreads();
        rz = srz;
        rv = srv;
        tz = stz;
        tv = stv;
    } else {
        rz = zd + (size_t)refr * nc;
        rv = img + (size_t)refr * nc;
        tz = zd + (size_t)tr * nc;
        tv = img + (size_t)tr * nc;
    }
    const float tr2 = (float)tr - xc, rr2 = (float)refr - xc;
    const int sz = 2 * maxdqz, sq = 2 * qmax;
    // column offsets -(qmax - 1) .. qmax - 1 over the threads, all reference columns per offset
    for (int oc = -(qmax - 1) + (int)threadIdx.x; oc < qmax; oc += blockDim.x) {
        const int c0 = max(0, -oc), c1 = min(nc, nc - oc);
        for (int rc = c0; rc < c1; ++rc) {
            const int tc = rc + oc;
            const float a = tr2 * tz[tc];
            const float b = rr2 * rz[rc];
            const int dqx = qmax + (int)rintf(a - b);
            const float ay = ((float)tc - yc) * tz[tc];
            const float by = ((float)rc - yc) * rz[rc];
            const int dqy = qmax + (int)rintf(ay - by);
            const int dqz = maxdqz + (int)rintf(cz * (rz[rc] - tz[tc]));
            if (dqx < 0 || dqx >= sq || dqy < 0 || dqy >= sq || dqz < 0 || dqz >= sz) continue;
            const double v = (double)rv[rc] * (double)tv[tc];
            unsafeAtomicAdd(out + ((size_t)dqz * sq + dqy) * sq + dqx, v);
        }
    }
}

// k_qcorr with a workgroup-private histogram in LDS.  For one (reference row, row
// offset) the bins' dqx index stays within a few values of qmax + offset (zd varies
// little along a row), so the workgroup accumulates all (dqz, dqy) bins of a window
// of `wx` dqx values starting at `lo` in LDS (FP64 LDS atomics) and flushes its
// non-zero bins to `out` with one global atomic each.  The window comes from the
// extremes of zd over the two rows; a pair whose bin falls outside it goes straight
// to `out`, so the result does not depend on that estimate.  Lanes take different
// column offsets (so their LDS atomics hit different bins) and all lanes are busy.
// Rows are staged in LDS (nc <= 1024).  Dynamic LDS: max(2 maxdqz x 2 qmax x wx
// doubles, 4 KB) for the histogram (first used for the zd extremes), then 4 nc floats.
__global__ __launch_bounds__(256) void k_qcorr_lds(const float* __restrict__ zd, const float* __restrict__ img,
                                                   int nr, int nc, float xc, float yc, float cz, int qmax,
                                                   int maxdqz, int wx, double* __restrict__ out)
{
#pragma clang fp contract(off)
    extern __shared__ double lds_d[];
    __shared__ int s_lo;
    const int refr = blockIdx.x, offr = blockIdx.y;
    const int tr = refr + offr;
    if (tr >= nr) return;  // whole workgroup: uniform
    const int sz = 2 * maxdqz, sq = 2 * qmax;
    const int nbins = sz * sq * wx;
    double* hist = lds_d;
    float(*red)[256] = reinterpret_cast<float(*)[256]>(lds_d);  // 4 x 256 floats, before hist is zeroed
    float* rows = reinterpret_cast<float*>(lds_d + (nbins > 512 ? nbins : 512));
    float* rz = rows;
    float* rv = rows + nc;
    float* tz = rows + 2 * nc;
    float* tv = rows + 3 * nc;
    float zmin_r = INFINITY, zmax_r = -INFINITY, zmin_t = INFINITY, zmax_t = -INFINITY;
    for (int i = threadIdx.x; i < nc; i += blockDim.x) {
        const float a = zd[(size_t)refr * nc + i], b = zd[(size_t)tr * nc + i];
        rz[i] = a;
        rv[i] = img[(size_t)refr * nc + i];
        tz[i] = b;
        tv[i] = img[(size_t)tr * nc + i];
        zmin_r = fminf(zmin_r, a);
        zmax_r = fmaxf(zmax_r, a);
        zmin_t = fminf(zmin_t, b);
        zmax_t = fmaxf(zmax_t, b);
    }
    red[0][threadIdx.x] = zmin_r;
    red[1][threadIdx.x] = zmax_r;
    red[2][threadIdx.x] = zmin_t;
    red[3][threadIdx.x] = zmax_t;
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {
        if ((int)threadIdx.x < s2) {
            red[0][threadIdx.x] = fminf(red[0][threadIdx.x], red[0][threadIdx.x + s2]);
            red[1][threadIdx.x] = fmaxf(red[1][threadIdx.x], red[1][threadIdx.x + s2]);
            red[2][threadIdx.x] = fminf(red[2][threadIdx.x], red[2][threadIdx.x + s2]);
            red[3][threadIdx.x] = fmaxf(red[3][threadIdx.x], red[3][threadIdx.x + s2]);
        }
        __syncthreads();
    }
    const float tr2 = (float)tr - xc, rr2 = (float)refr - xc;
    if (threadIdx.x == 0) {
        // tr2 zd_t - rr2 zd_r is linear in each zd: its extremes are at the corners
        float vmin = INFINITY;
        const float zts[2] = {red[2][0], red[3][0]}, zrs[2] = {red[0][0], red[1][0]};
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) vmin = fminf(vmin, tr2 * zts[i] - rr2 * zrs[j]);
        s_lo = qmax + (int)rintf(vmin);
    }
    __syncthreads();  // red has been read: the space becomes the histogram
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) hist[b] = 0.0;
    __syncthreads();
    const int lo = s_lo;
    const int noc = 2 * qmax - 1;
    const int64_t nitems = (int64_t)noc * nc;
    for (int64_t it = threadIdx.x; it < nitems; it += blockDim.x) {
        const int oc = (int)(it % noc) - (qmax - 1);
        const int rc = (int)(it / noc);
        const int tc = rc + oc;
        if (tc < 0 || tc >= nc) continue;
        const float a = tr2 * tz[tc];
        const float b = rr2 * rz[rc];
        const int dqx = qmax + (int)rintf(a - b);
        const float ay = ((float)tc - yc) * tz[tc];
        const float by = ((float)rc - yc) * rz[rc];
        const int dqy = qmax + (int)rintf(ay - by);
        const int dqz = maxdqz + (int)rintf(cz * (rz[rc] - tz[tc]));
        if (dqx < 0 || dqx >= sq || dqy < 0 || dqy >= sq || dqz < 0 || dqz >= sz) continue;
        const double v = (double)rv[rc] * (double)tv[tc];
        const int wi = dqx - lo;
        if (wi >= 0 && wi < wx)
            atomicAdd(hist + ((size_t)dqz * sq + dqy) * wx + wi, v);
        else
            unsafeAtomicAdd(out + ((size_t)dqz * sq + dqy) * sq + dqx, v);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbins; b += blockDim.x) {
        const double v = hist[b];
        if (v != 0.0) {
            const int wi = b % wx;
            const int r = b / wx;
            const int dqx = lo + wi;
            if (dqx >= 0 && dqx < sq) unsafeAtomicAdd(out + (size_t)r * sq + dqx, v);
        }
    }
}

// One thread per bin with dqx index < qmax; the mirror has dqx index (sq - i2) mod sq,
// which is >= qmax except for i2 == 0 (its own column): there only the lower of the
// two flat indices writes, so every pair is updated exactly once.
__global__ __launch_bounds__(256) void k_qcorr_assemble(double* __restrict__ out, int maxdqz, int qmax)
{
    const int sz = 2 * maxdqz, sq = 2 * qmax;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nh = (int64_t)sz * sq * qmax;
    if (idx >= nh) return;
    const int i2 = (int)(idx % qmax);
    const int64_t r = idx / qmax;
    const int i1 = (int)(r % sq), i0 = (int)(r / sq);
    const int m0 = (sz - i0) % sz, m1 = (sq - i1) % sq, m2 = (sq - i2) % sq;
    const size_t a = ((size_t)i0 * sq + i1) * sq + i2;
    const size_t b = ((size_t)m0 * sq + m1) * sq + m2;
    if (a == b) {  // its own mirror (dqx index 0 only): added to itself, as in the reference
        out[a] = out[a] + out[a];
        return;
    }
    if (i2 == 0 && b < a) return;  // the pair is handled by its other member
    const double s = out[a] + out[b];
    out[a] = s;
    out[b] = s;
}

hipError_t launch_corr_accumulate(const float* corr, const float* maskcorr, int64_t n, int64_t count, float* mean,
                                  float* m2, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_corr_accumulate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, corr, maskcorr, n,
                       count, mean, m2);
    return hipGetLastError();
}

hipError_t launch_qcorr(const float* zd, const float* img, int nr, int nc, float xc, float yc, float cz, int qmax,
                        int maxdqz, double* out, hipStream_t st)
{
    const int64_t nout = (int64_t)(2 * maxdqz) * (2 * qmax) * (2 * qmax);
    hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, st, out, nout, 0.0);
    // LDS histogram window: as many dqx values as fit with the rows in 78 KB (two workgroups per CU)
    const size_t row_bytes = (size_t)4 * nc * sizeof(float);
    const size_t per_dqx = (size_t)(2 * maxdqz) * (2 * qmax) * sizeof(double);
    const size_t budget = (size_t)78 * 1024;
    const int wx = row_bytes < budget ? (int)std::min<size_t>(8, (budget - row_bytes) / per_dqx) : 0;
    if (nc <= 1024 && wx >= 2) {
        const size_t lds = std::max<size_t>((size_t)wx * per_dqx, 4096) + row_bytes;
        hipLaunchKernelGGL(k_qcorr_lds, dim3((unsigned)nr, (unsigned)qmax), dim3(256), lds, st, zd, img, nr, nc, xc,
                           yc, cz, qmax, maxdqz, wx, out);
    } else {
        const size_t lds = nc <= 1024 ? row_bytes : 0;
        hipLaunchKernelGGL(k_qcorr, dim3((unsigned)nr, (unsigned)qmax), dim3(256), lds, st, zd, img, nr, nc, xc, yc,
                           cz, qmax, maxdqz, out);
    }
    const int64_t nh = (int64_t)(2 * maxdqz) * (2 * qmax) * qmax;
    hipLaunchKernelGGL(k_qcorr_assemble, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, st, out, maxdqz, qmax);
    return hipGetLastError();
}

}  // namespace idisim
