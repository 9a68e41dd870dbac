// Launchers of the device kernels (kernels.hip, adjoint.hip); used by the C ABI in idisim.hip.
#pragma once
#include "geometry.hpp"

namespace idisim {

constexpr int OUT_FIELD = 0;
constexpr int OUT_INTENSITY = 1;

// FP32-split forward.  v: tile variant, w: curvature/second-order term present,
// a: per-emitter amplitude present, s: sum mode (0 chunk+Kahan, 1 Kahan per term).
// splits > 1 writes float4 partials [splits][n0*n1] to `partial`.
hipError_t launch_fwd_split(int v, bool w, bool a, int s, const Geom& g, const double* emit, const double* amp,
                            int64_t n, int splits, float4* partial, void* out, int kind, hipStream_t st);
hipError_t launch_finalize_split(const Geom& g, const float4* partial, int splits, void* out, int kind,
                                 hipStream_t st);
hipError_t launch_batched_split(int v, bool w, bool a, int s, const Geom& g, const double* emit,
                                const double* amp, const WorkItem* items, int nitems, int* counter, int nblocks,
                                void* out, int kind, hipStream_t st);
int occupancy_batched_split(int v, bool w, bool a, int s);

// FP32-split forward with the phase on the matrix cores (tile variant kTileMfma).
hipError_t launch_fwd_mfma(const Geom& g, const double* emit, const double* amp, int64_t n, int splits,
                           float4* partial, void* out, int kind, hipStream_t st);
hipError_t launch_batched_mfma(const Geom& g, const double* emit, const double* amp, const WorkItem* items,
                               int nitems, int* counter, int nblocks, void* out, int kind, hipStream_t st);
int occupancy_batched_mfma(bool fold);

// Literal FP64 forward.
hipError_t launch_fwd_fp64(int v, const Geom& g, const double* emit, const double* amp, int64_t n, int splits,
                           double4* partial, void* out, int kind, hipStream_t st);
hipError_t launch_finalize_fp64(const Geom& g, const double4* partial, int splits, void* out, int kind,
                                hipStream_t st);
hipError_t launch_batched_fp64(int v, const Geom& g, const double* emit, const double* amp,
                               const WorkItem* items, int nitems, int* counter, int nblocks, void* out, int kind,
                               hipStream_t st);
int occupancy_batched_fp64(int v);

// Adjoint (FP32-split).  pixel_splits > 1 writes double partials
// [pixel_splits][n][5] to `partial` and a finalize kernel sums them.
hipError_t launch_adjoint(bool w, const Geom& g, const double* emit, const double* amp, int64_t n,
                          const float2* field, const float* dLdI, int pixel_splits, double* partial,
                          float* grad_pos, float* grad_phase, float* grad_amp, hipStream_t st);

hipError_t launch_extent(const double* emit, int64_t n, unsigned long long* out, hipStream_t st);

// Emitters in a ball / random phases from Philox4x32-10 (generate.hip).
hipError_t launch_gen_ball(uint64_t seed, uint64_t stream, int64_t n, double radius, const double* centre,
                           int positions, int phases, double* out, hipStream_t st);

}  // namespace idisim
