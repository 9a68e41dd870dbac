// Launchers of the device kernels (kernels.hip, adjoint.hip); used by the C ABI in idisim.hip.
#pragma once
#include "geometry.hpp"

namespace idisim {

// The Geom argument of a kernel, read again from the kernarg segment (a kernel's first argument
// sits at offset 0 of that segment).  The persistent kernels read it for each work item: held as a
// by-value argument, all of it stays live in SGPRs across the emitter loop of every item, and the
// excess is spilled into VGPR lanes (32 SGPRs and 28 VGPRs in the mirror-pair kernels, one wave
// per SIMD fewer).  A second read after a long loop also keeps the compiler from holding FP64
// values formed before the loop (pixel frames) in VGPRs across it, when the code after the loop
// forms the same values again.  The empty asm makes the pointer opaque, so the loads are not
// hoisted; they are scalar loads from constant memory.
__device__ __forceinline__ Geom item_geom()
{
    static_assert(sizeof(Geom) % 4 == 0, "Geom is copied as 32-bit words");
    using KernargWords = const __attribute__((address_space(4))) uint32_t*;
    KernargWords p = (KernargWords)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(p));
    Geom g;
    uint32_t* d = reinterpret_cast<uint32_t*>(&g);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(Geom) / 4); ++i) d[i] = p[i];
    return g;
}

// Wave priority of a pixel or emitter loop from the share of its work that a wave has done, q of
// 32 (s_setprio; the kernels run their FP64 record stages at 3, above every loop).  Waves of
// different workgroups that share a SIMD otherwise issue oldest first: the oldest workgroup runs
// ahead, and the last one finishes alone on its SIMD at a fraction of the issue rate (measured at
// 4096 emitters: the four workgroups of a CU ended 36, 47, 63 and 78 us into the adjoint).  With
// the waves that are behind issuing first, they finish together; the last 1/16 of the work has a
// level of its own.
__device__ __forceinline__ void prio_by_progress(int q)
{
    if (q < 26)
        __builtin_amdgcn_s_setprio(2);
    else if (q < 30)
        __builtin_amdgcn_s_setprio(1);
    else
        __builtin_amdgcn_s_setprio(0);
}

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
// Mirror-pair tiles (tile_pair(v)); W must be folded into A (g.foldw) when the
// phase model has a curvature / second-order term.  tab: the variant's set-up table
// (launch_pair_table) or nullptr, when the kernel forms it in its prologue.
hipError_t launch_fwd_pair(int v, bool a, const Geom& g, const double* emit, const double* amp, int64_t n,
                           int splits, const float* tab, float4* partial, void* out, int kind, hipStream_t st);
// The set-up table of a mirror-pair variant with one quad per lane (0 bytes: the variant has none):
// the per-lane quad terms, reference data and tile-centre frame, from the geometry alone.
size_t pair_table_bytes(int v, const Geom& g);
hipError_t launch_pair_table(int v, const Geom& g, float* tab, hipStream_t st);
hipError_t launch_batched_pair(int v, bool a, const Geom& g, const double* emit, const double* amp,
                               const WorkItem* items, int nitems, int* counter, int nblocks, void* out, int kind,
                               hipStream_t st);
int occupancy_batched_pair(int v, bool a);

// Literal FP64 forward.
hipError_t launch_fwd_fp64(int v, const Geom& g, const double* emit, const double* amp, int64_t n, int splits,
                           double4* partial, void* out, int kind, hipStream_t st);
hipError_t launch_finalize_fp64(const Geom& g, const double4* partial, int splits, void* out, int kind,
                                hipStream_t st);
hipError_t launch_batched_fp64(int v, const Geom& g, const double* emit, const double* amp,
                               const WorkItem* items, int nitems, int* counter, int nblocks, void* out, int kind,
                               hipStream_t st);
int occupancy_batched_fp64(int v);

// Adjoint (FP32-split) over 16 x 16 pixel tiles.  The scratch buffer holds the per-pixel
// tile data (float4 + float2 per pixel) followed by double partials [splits][n][5].
struct AdjointLayout {
    int tiles_y, ntiles, per, splits;  // 16 x 16 tiles along y, in all, per workgroup; splits
    int64_t egroups;
    size_t pix_bytes, part_bytes, bytes;  // pixel tables, split partials, all (with ntiles loss parts)
};
AdjointLayout adjoint_layout(const Geom& g, int64_t n, int num_cus, int blocks_per_cu);
int adjoint_occupancy(bool w, bool lam, bool amp);  // resident workgroups per CU
// w: lam_p W_cj enters the phase (not folded); lam: far-field secondorder with w.
// target == nullptr: dL/dI is dLdI; else dL/dI = 2 (|E|^2 - target), and loss (device, may be
// NULL) receives sum (|E|^2 - target)^2.  fwd_partial (with a target): the field is not read
// from `field` but summed from a mirror-pair forward's float4 split partials [fwd_splits][npix],
// as the forward's finalize would, and stored to field_out when that is not NULL.
hipError_t launch_adjoint(bool w, bool lam, const Geom& g, const AdjointLayout& L, const double* emit,
                          const double* amp, int64_t n, const float2* field, const float* dLdI, const float* target,
                          double* loss, void* scratch, float* grad_pos, float* grad_phase, float* grad_amp,
                          hipStream_t st, const float4* fwd_partial = nullptr, int fwd_splits = 0,
                          float2* field_out = nullptr);

hipError_t launch_extent(const double* emit, int64_t n, unsigned long long* out, hipStream_t st);
// issue-rate probe (kind 0: v_fma_f32, 1: v_sin_f32 / v_cos_f32); out: blocks x 256 floats
hipError_t launch_issue_probe(int kind, int blocks, int iters, float* out, hipStream_t st);
hipError_t launch_zero(void* out, size_t bytes, hipStream_t st);  // bytes: a multiple of 4

// Emitters in a ball / random phases from Philox4x32-10 (generate.hip).
hipError_t launch_gen_ball(uint64_t seed, uint64_t stream, int64_t n, double radius, const double* centre,
                           int positions, int phases, double* out, hipStream_t st);

// Emitters in several balls (offsets: CSR over balls, device) and in a Gaussian cloud.
struct Rot3 {
    double m[9];
};
hipError_t launch_gen_balls(uint64_t seed, uint64_t stream, int64_t n, double radius, int64_t nballs,
                            const double* centres, const int64_t* offsets, int positions, int phases, double* out,
                            hipStream_t st);
hipError_t launch_gen_gauss(uint64_t seed, uint64_t stream, int64_t n, const double* sigma, const double* rot,
                            int positions, int phases, double* out, hipStream_t st);

// Correlators (correlate.hip).
hipError_t launch_corr_accumulate(const float* corr, const float* maskcorr, int64_t n, int64_t count, float* mean,
                                  float* m2, hipStream_t st);
hipError_t launch_qcorr(const float* zd, const float* img, int nr, int nc, float xc, float yc, float cz, int qmax,
                        int maxdqz, double* out, hipStream_t st);

// Time-dependent simulation (timesim.hip): pixels are processed in batches of np; the scratch
// holds keys (2x), indices (2x), amplitudes, segment offsets and the sort's temporary storage.
struct TimeLayout {
    int np = 1;
    bool fp64 = false;
    size_t keys_bytes = 0, idx_bytes = 0, amp_bytes = 0, off_bytes = 0, sort_bytes = 0, bytes = 0;
};
TimeLayout time_layout(const Geom& g, int64_t n, bool fp64_keys, size_t budget);
// path: 0 = by emitter count, 1 = one workgroup per pixel (n <= 10240), 2 = batched sort path
hipError_t launch_time(const Geom& g, const TimeLayout& L, const double* emit, const double* times, int64_t n,
                       double c, double tau, int path, void* scratch, double* out, hipStream_t st);
bool time_fused_ok(int64_t n, bool fp64_keys, int path);

// Choice of n of m sites without replacement, optionally weighted (choice.hip).
size_t choice_workspace_bytes(int64_t m);
hipError_t launch_choice(uint64_t seed, uint64_t stream, int64_t n, int64_t m, const double* sites, const double* w,
                         const double* rot, int phases, double* out, void* ws, hipStream_t st);

// Detector charge sharing / per-pixel gain noise (photons.hip).
hipError_t launch_chargesharing(const double* counts, int n0, int n1, double s, double ns, uint64_t seed,
                                double* out, hipStream_t st);

}  // namespace idisim
