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
    const size_t lds = nc <= 1024 ? (size_t)4 * nc * sizeof(float) : 0;
    hipLaunchKernelGGL(k_qcorr, dim3((unsigned)nr, (unsigned)qmax), dim3(256), lds, st, zd, img, nr, nc, xc, yc, cz,
                       qmax, maxdqz, out);
    const int64_t nh = (int64_t)(2 * maxdqz) * (2 * qmax) * qmax;
    hipLaunchKernelGGL(k_qcorr_assemble, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, st, out, maxdqz, qmax);
    return hipGetLastError();
}

}  // namespace idisim
