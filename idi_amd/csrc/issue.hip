// Issue-rate probes for the instruction kinds that bound the FP32-split kernels:
// v_fma_f32 and v_sin_f32 / v_cos_f32.  Every lane runs eight independent chains of
// one kind (no dependency stalls with 8 waves per SIMD), unrolled so that the loop's
// scalar work is a small part of the issue; the host times a launch and
// divides the lane operations by the time.  No counterpart in the reference: they
// state the issue ceiling that the forward kernels are measured against.
#include "kernels.hpp"

namespace idisim {

namespace {
// a_k += m_k q_k in the accumulator form v_fmac_f32 (32-bit encoding, every operand in a
// VGPR), 64 per loop iteration so that the loop's scalar work stays off the count.  The
// additions cannot be reassociated (no fast-math), so none is folded away.
__global__ __launch_bounds__(256) void k_issue_fma(float* __restrict__ out, int iters, float seed)
{
    float a[8], m[8], q[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        a[k] = seed + 1e-3f * (float)(threadIdx.x + k);
        m[k] = 1e-3f * (float)(1 + ((threadIdx.x + 3 * k) & 7));
        q[k] = 1e-4f * (float)(1 + ((threadIdx.x + 5 * k) & 15));
    }
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
#pragma unroll
            for (int k = 0; k < 8; ++k) a[k] = fmaf(m[k], q[k], a[k]);
        }
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) s += a[k];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// eight chains of v_sin_f32 / v_cos_f32, 32 per loop iteration
__global__ __launch_bounds__(256) void k_issue_trans(float* __restrict__ out, int iters, float seed)
{
    float a[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = seed + 1e-3f * (float)(threadIdx.x + k);
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
#pragma unroll
            for (int k = 0; k < 8; k += 2) {
                a[k] = __builtin_amdgcn_sinf(a[k]);
                a[k + 1] = __builtin_amdgcn_cosf(a[k + 1]);
            }
        }
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) s += a[k];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}
}  // namespace

// out holds blocks x 256 floats; 64 (kind 0) or 32 (kind 1) lane operations per lane and iteration.
hipError_t launch_issue_probe(int kind, int blocks, int iters, float* out, hipStream_t st)
{
    if (kind == 0)
        hipLaunchKernelGGL(k_issue_fma, dim3(blocks), dim3(256), 0, st, out, iters, 0.25f);
    else
        hipLaunchKernelGGL(k_issue_trans, dim3(blocks), dim3(256), 0, st, out, iters, 0.25f);
    return hipGetLastError();
}

}  // namespace idisim
