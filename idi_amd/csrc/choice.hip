// On-device choice of N lattice sites without replacement (no reference counterpart; it moves
// the per-image draw of simobj.crystal, simobj.py:479-487, onto the GPU).
//
// Weighted sampling without replacement by exponential keys (Efraimidis and Spirakis,
// "Weighted random sampling with a reservoir", Inf. Proc. Letters 97 (2006)): site i gets
// key -ln(u_i) / w_i with u_i uniform in (0, 1]; the N smallest keys are a sample with the
// distribution of N successive draws proportional to the remaining weights, which is what
// numpy's choice(replace=False, p=w) draws.  Unit weights give a uniform N-subset.  More
// sites than exist are taken in rounds, each round a fresh draw of all sites (as the
// reference's loop over choice calls does).
//
// Random numbers (Philox4x32-10, key = seed): site i in round r uses block
// (i_lo, i_hi, stream, 0x100 + r); emitter j's phase uses block (j_lo, j_hi, stream, 0x200).
// The keys are sorted with rocPRIM's radix sort (stable), the chosen sites are gathered in
// key order and rotated by `rot` when given.
#include <cstring>  // host memset, used by rocPRIM headers

#include <rocprim/rocprim.hpp>

#include "kernels.hpp"
#include "philox.hpp"

namespace idisim {

__global__ void k_choice_keys(uint64_t seed, uint64_t stream, int round, int64_t m, const double* __restrict__ w,
                              double* __restrict__ keys, int* __restrict__ idx)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const U4 c = {{(uint32_t)i, (uint32_t)((uint64_t)i >> 32), (uint32_t)stream, 0x100u + (uint32_t)round}};
    const U4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double u = 1.0 - u53(r.v[0], r.v[1]);  // (0, 1]
    const double wi = (w != nullptr) ? w[i] : 1.0;
    keys[i] = (wi > 0.0) ? -log(u) / wi : INFINITY;
    idx[i] = (int)i;
}

__global__ void k_choice_gather(const int* __restrict__ idx, int64_t take, int64_t done,
                                const double* __restrict__ sites, Rot3 M, int rotate, double* __restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= take) return;
    const double* s = sites + 3 * (int64_t)idx[t];
    double* o = out + 4 * (done + t);
    if (rotate) {
        o[0] = M.m[0] * s[0] + M.m[1] * s[1] + M.m[2] * s[2];
        o[1] = M.m[3] * s[0] + M.m[4] * s[1] + M.m[5] * s[2];
        o[2] = M.m[6] * s[0] + M.m[7] * s[1] + M.m[8] * s[2];
    } else {
        o[0] = s[0];
        o[1] = s[1];
        o[2] = s[2];
    }
}

__global__ void k_choice_phases(uint64_t seed, uint64_t stream, int64_t n, double* __restrict__ out)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const U4 c = {{(uint32_t)j, (uint32_t)((uint64_t)j >> 32), (uint32_t)stream, 0x200u}};
    const U4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    out[4 * j + 3] = kTwoPi * u53(r.v[0], r.v[1]);
}

static size_t a256(size_t b) { return (b + 255) / 256 * 256; }

size_t choice_workspace_bytes(int64_t m)
{
    size_t sort = 0;
    (void)rocprim::radix_sort_pairs((void*)nullptr, sort, (const double*)nullptr, (double*)nullptr,
                                    (const int*)nullptr, (int*)nullptr, (unsigned)m);
    return 2 * a256((size_t)m * sizeof(double)) + 2 * a256((size_t)m * sizeof(int)) + a256(sort);
}

hipError_t launch_choice(uint64_t seed, uint64_t stream, int64_t n, int64_t m, const double* sites, const double* w,
                         const double* rot, int phases, double* out, void* ws, hipStream_t st)
{
    char* base = static_cast<char*>(ws);
    const size_t kb = a256((size_t)m * sizeof(double)), ib = a256((size_t)m * sizeof(int));
    double* k0 = reinterpret_cast<double*>(base);
    double* k1 = reinterpret_cast<double*>(base + kb);
    int* i0 = reinterpret_cast<int*>(base + 2 * kb);
    int* i1 = reinterpret_cast<int*>(base + 2 * kb + ib);
    void* tmp = base + 2 * kb + 2 * ib;
    size_t tmp_bytes = choice_workspace_bytes(m) - (2 * kb + 2 * ib);
    Rot3 M{};
    if (rot)
        for (int k = 0; k < 9; ++k) M.m[k] = rot[k];
    const unsigned gm = (unsigned)((m + 255) / 256);
    int64_t done = 0;
    for (int round = 0; done < n; ++round) {
        const int64_t take = (n - done) < m ? (n - done) : m;
        hipLaunchKernelGGL(k_choice_keys, dim3(gm), dim3(256), 0, st, seed, stream, round, m, w, k0, i0);
        hipError_t e = rocprim::radix_sort_pairs(tmp, tmp_bytes, k0, k1, i0, i1, (unsigned)m, 0, 64, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_choice_gather, dim3((unsigned)((take + 255) / 256)), dim3(256), 0, st, i1, take, done,
                           sites, M, rot ? 1 : 0, out);
        done += take;
    }
    if (phases && n > 0)
        hipLaunchKernelGGL(k_choice_phases, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, seed, stream, n, out);
    return hipGetLastError();
}

}  // namespace idisim
