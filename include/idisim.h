/*
 * idisim -- MI355X (gfx950) direct-sum wavefield simulation, C ABI.
 *
 * For every detector pixel p the library evaluates
 *     E_p = sum_j a_j m_pj exp(i theta_pj),   I_p = |E_p|^2
 * with the phase models of idi's simulation package:
 *   near field ("nf"):  theta = k (|r_p - r_j| - z_p + z_j) + phi_j
 *                       (idi/simulation/cpu.py:42-55, sim.cu:43-69)
 *   far field:          theta = -q_p . r_j + phi_j,  q_p = k (r_p/|r_p| - z_hat)
 *                       (cpu.py:57-66,85-89, sim.cu:71-95), optional
 *                       "secondorder" + k/(2|r_p|) (x_j^2 + y_j^2) (sim.cu:88)
 *   m_pj = 1, or 1/|r_p - r_j| (nf) / 1/|r_p| (ff) with "scale"
 *   (cpu.py:48-54,60-65).
 * Pixel (x, y), x the row index: r_p = M ((x - n0/2) s + off0, (y - n1/2) s + off1, z)
 * with true division (cpu.py:76-80) and M the optional detector rotation.
 *
 * Conventions
 *   - Every array argument is a DEVICE pointer unless its comment says host.
 *     The caller owns all buffers; the library never frees caller memory.
 *   - Emitters are the reference layout: n x 4 row-major float64 (x, y, z, phi),
 *     as simobj.get() returns them (simobj.py:63-66, cpu.py:69).
 *   - Return value 0 = success, < 0 = error; idisim_last_error() gives a
 *     thread-local message.  The library never calls exit() or abort().
 *   - A handle is used by one host thread / stream at a time; different
 *     handles are independent.  Work is enqueued on `stream` (a hipStream_t;
 *     NULL = the legacy default stream) and is asynchronous unless stated.
 */
#ifndef IDISIM_H
#define IDISIM_H

#include <stddef.h>
#include <stdint.h>

/* The library is built with hidden visibility; only these entry points are exported. */
#if defined(__GNUC__) || defined(__clang__)
#define IDISIM_API __attribute__((visibility("default")))
#else
#define IDISIM_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define IDISIM_VERSION 2

/* geometry flags (settings substrings of cuda.py:53-74 / cpu.py:189-190) */
#define IDISIM_NF          1u  /* "nf": exact distance (near field)            */
#define IDISIM_SCALE       2u  /* "scale": 1/r amplitude                       */
#define IDISIM_SECONDORDER 4u  /* "secondorder": far-field Fresnel term       */
#define IDISIM_ROTATE      8u  /* apply rot[] to the pixel positions          */

/* precision of a handle */
#define IDISIM_PREC_FP64      0 /* literal FP64 formula, complex128 output      */
#define IDISIM_PREC_FP32SPLIT 1 /* FP64 range-reduced phase, FP32 sincos/sums,
                                   complex64 output                            */

/* output kinds */
#define IDISIM_OUT_FIELD     0 /* complex (re, im) interleaved, n0 x n1        */
#define IDISIM_OUT_INTENSITY 1 /* |E|^2, n0 x n1                               */

/* error codes */
#define IDISIM_OK            0
#define IDISIM_E_INVALID    -1 /* invalid argument                             */
#define IDISIM_E_HIP        -2 /* HIP runtime error                            */
#define IDISIM_E_NOMEM      -3 /* device allocation failed                     */
#define IDISIM_E_PRECISION  -4 /* FP32-split error bound not met (see options) */
#define IDISIM_E_UNSUPPORTED -5

/* options (idisim_set_option) */
#define IDISIM_OPT_AUTOTUNE     1 /* 1 (default): time candidate plans on first use */
#define IDISIM_OPT_TILE         2 /* force a tile variant (-1 = planner): 0..3 VALU
                                     (16x16 .. 32x64 px), 4 reserved (no kernel),
                                     5, 6 mirror pairs (FP32-split only, 32x32 and
                                     32x64 px), 7 mirror pairs with a centre per
                                     pair (FP32-split only, 32x32 px), 8 four
                                     mirror pairs per 2x4 block (FP32-split only,
                                     32x64 px), 9, 10 (as 5) and 11, 12 (as 7)
                                     with the pair phasors rotated from a
                                     per-wave reference to first (9, 11) or
                                     second (10, 12) order, 13 as 9 with the
                                     rotation angle linear in the quad offset,
                                     14 as 11 with a reference per 8x8 px
                                     sub-block, 15 as 14 with the rotation
                                     between the pair centres also from the
                                     sub-block reference (FP32-split only,
                                     32x32 px);
                                     FP64 handles take 0..3                      */
#define IDISIM_OPT_SPLIT        3 /* force emitter split count (0 = planner)        */
#define IDISIM_OPT_SUMMODE      4 /* 0 (default): per-chunk partials merged with
                                     Kahan; 1: Kahan on every term (cpu.py:97-105) */
#define IDISIM_OPT_GRAPH        5 /* 1 (default): batched calls replay a hipGraph   */
#define IDISIM_OPT_TOL_PPB      6 /* per-term phase error bound, parts per 1e9 of a
                                     revolution (default 5000 = 5e-6 rev)          */
#define IDISIM_OPT_FALLBACK     7 /* 1 (default): FP32-split handles run the FP64
                                     kernel when no tile meets the bound; 0: error */
#define IDISIM_OPT_TIME_PATH    8 /* time-dependent simulation: 0 (default) and 1
                                     take one workgroup per pixel when it fits
                                     (n <= 20480 with float32 keys, 12288 with
                                     float64 keys), else the batched sort path;
                                     2 = always the batched sort path          */

typedef struct idisim_handle idisim_handle;

typedef struct idisim_geometry {
    int64_t n0, n1;       /* detector pixels (rows, cols)                  */
    double pixelsize;     /* pixel pitch, same length unit as emitters     */
    double detz;          /* detector distance                             */
    double k;             /* angular wavenumber (1 / length unit)          */
    double detoffset[2];  /* detector centre offset (x, y)                 */
    double rot[9];        /* row-major rotation (IDISIM_ROTATE)            */
    uint32_t flags;       /* IDISIM_NF | IDISIM_SCALE | ...                */
} idisim_geometry;

/* Plan actually used by the last forward/adjoint call of a handle. */
typedef struct idisim_plan {
    int32_t precision;    /* precision that ran (FP64 after a fallback)    */
    int32_t tile;         /* tile variant index                            */
    int32_t tile_rows, tile_cols; /* pixels per workgroup tile             */
    int32_t split;        /* emitter split count                           */
    int32_t summode;
    int32_t graph;        /* 1 if a hipGraph was replayed                  */
    int32_t fallback;     /* 1 if FP32-split fell back to FP64             */
    double phase_bound;   /* a-priori per-term phase error bound, rev      */
    double tuned_ms;      /* best candidate time measured by the autotuner */
    int32_t extent_exceeded; /* 1: a finished FP32-split kernel of this handle met an
                                emitter with |r_j| > the extent of its call, so its
                                phase_bound did not hold (read after synchronising;
                                the next call then returns IDISIM_E_PRECISION) */
    int32_t reserved;
} idisim_plan;

/* Size of the static device scratch of idisim_forward_noalloc (per device). */
#define IDISIM_NOALLOC_SCRATCH_BYTES (64u << 20)

/* Library ABI version (IDISIM_VERSION). */
IDISIM_API int idisim_version(void);

/* Thread-local message of the last error ("" if none). */
IDISIM_API const char *idisim_last_error(void);

/* Create a handle on `device` for a fixed geometry and precision.
   Replaces the per-call kernel set-up of cuda.py:53-86 (flags, rotation,
   module compile) and the closure of cpu.py:6-40 (get_kernel). */
IDISIM_API int idisim_create(idisim_handle **out, const idisim_geometry *geom, int precision, int device);

/* Free the handle's device buffers and graph (no reference counterpart: cupy
   frees the generator's buffers when they are garbage-collected). */
IDISIM_API int idisim_destroy(idisim_handle *h);

/* Set an IDISIM_OPT_* option; clears the handle's autotune cache (no
   reference counterpart). */
IDISIM_API int idisim_set_option(idisim_handle *h, int option, int64_t value);

/* The plan of the handle's last forward / batched call (no reference
   counterpart). */
IDISIM_API int idisim_get_plan(const idisim_handle *h, idisim_plan *out);

/* One shot.  Replaces the kernels wfkernelf / wfkerneld (sim.cu:141-149) as
   launched by cuda.py:91-107, and the numba kernel(ret, atom) of cpu.py:68-107.
     emitters:   n x 4 float64 (x, y, z, phi)
     amplitudes: n float64 or NULL (all ones; the reference has no amplitudes)
     extent:     an upper bound of max_j |r_j| (used by the FP32-split error
                 planner); <= 0 lets the library measure it (synchronises).  The
                 FP32-split kernels check it: an emitter beyond it sets a flag, and
                 the handle's next forward / batched / adjoint call returns
                 IDISIM_E_PRECISION (idisim_get_plan reports it as extent_exceeded)
     out:        complex64 (FP32SPLIT) / complex128 (FP64) n0 x n1 for FIELD,
                 float32 / float64 n0 x n1 for INTENSITY, row-major, ret[x, y] */
IDISIM_API int idisim_forward(idisim_handle *h, const double *emitters, const double *amplitudes,
                   int64_t n, double extent, void *out, int out_kind, void *stream);

/* idisim_forward without device allocation.  No counterpart in the reference: for
   callers that cannot call hipMalloc (code under HIP graph capture, serving loops,
   harnesses that forbid allocation).  The emitter-split partial sums (and the extent
   word when extent <= 0) live in a static device buffer of
   IDISIM_NOALLOC_SCRATCH_BYTES that each device holds for the life of the process;
   the split count is capped so that its partials fit (n0 n1 x 16 B per split on
   FP32-split handles, 32 B on FP64), and a call that needs more even at one split
   (an FP64 fallback on a detector above 2^21 pixels) fails with IDISIM_E_INVALID.
   The plan is the handle's tuned one (from earlier idisim_forward calls) or the
   planner's heuristic; no timing pass runs.  Every call on a device shares that
   device's buffer: calls that overlap in time on one device (different streams,
   host threads or handles) race.  extent <= 0 is measured through the buffer and
   synchronises `stream` (not allowed under graph capture: pass extent > 0). */
IDISIM_API int idisim_forward_noalloc(idisim_handle *h, const double *emitters, const double *amplitudes,
                                      int64_t n, double extent, void *out, int out_kind, void *stream);

/* Many shots with uneven emitter counts in one persistent launch (replayed as
   a hipGraph when the buffers and counts repeat).  No reference counterpart;
   it batches the per-image loop of cuda.py:108-136 / cpu.py:207-212.
     emitters:   concatenated (sum n_s) x 4 float64
     offsets:    HOST array of nshots + 1 int64, offsets[0] = 0
     out:        nshots x n0 x n1 of the output kind/type of idisim_forward */
IDISIM_API int idisim_forward_batched(idisim_handle *h, const double *emitters, const double *amplitudes,
                           const int64_t *offsets, int64_t nshots, double extent,
                           void *out, int out_kind, void *stream);

/* Adjoint of I = |E|^2 for a loss L(I) (no reference counterpart).
     field:      the complex64 field that idisim_forward returned for the same
                 emitters (FP32SPLIT handles only).  With IDISIM_SCALE the
                 gradient of the 1/r amplitude itself is left out (smaller than
                 the phase term by 1/(k r), ~1e-9 for X-rays)
     dLdI:       float32 n0 x n1
     grad_pos:   float32 n x 3  dL/d(x, y, z)
     grad_phase: float32 n      dL/dphi
     grad_amp:   float32 n      dL/da (may be NULL)                        */
IDISIM_API int idisim_adjoint(idisim_handle *h, const double *emitters, const double *amplitudes,
                   int64_t n, double extent, const void *field, const float *dLdI,
                   float *grad_pos, float *grad_phase, float *grad_amp, void *stream);

/* idisim_adjoint for the squared-intensity loss L = sum_p (|E_p|^2 - T_p)^2 (no reference
   counterpart; the loss of the fitting workload): dL/dI = 2 (|E|^2 - T) is formed from `field`
   and `target` inside the adjoint's pixel pass, so the caller runs no kernels of its own.
     target: float32 n0 x n1
     loss:   one float64, DEVICE, or NULL: receives L (summed per 16 x 16 tile, then over the
             tiles in a fixed order, so it is the same bit for bit on every call)
   The other arguments are those of idisim_adjoint; n >= 1 when loss is requested. */
IDISIM_API int idisim_adjoint_l2(idisim_handle *h, const double *emitters, const double *amplitudes,
                                 int64_t n, double extent, const void *field, const float *target, double *loss,
                                 float *grad_pos, float *grad_phase, float *grad_amp, void *stream);

/* The squared-intensity loss and its gradients in one call (no reference counterpart; a fitting
   step): idisim_forward of the emitters (IDISIM_OUT_FIELD) followed by idisim_adjoint_l2 on its
   field, with the same results bit for bit.  When the forward's plan is a mirror-pair tile with
   emitter splits, the adjoint's pixel pass sums the forward's split partials itself: one launch
   where the forward's finalize and the pixel pass took two.
     field: complex64 n0 x n1, DEVICE, or NULL: receives the forward's field
   The other arguments are those of idisim_adjoint_l2 (target and loss as there); n >= 1. */
IDISIM_API int idisim_loss_grad_l2(idisim_handle *h, const double *emitters, const double *amplitudes,
                                   int64_t n, double extent, const float *target, void *field, double *loss,
                                   float *grad_pos, float *grad_phase, float *grad_amp, void *stream);

/* Largest |r_j| of n emitters (device reduction; synchronises `stream`). */
IDISIM_API int idisim_extent(const double *emitters, int64_t n, double *extent_out, void *stream);

/* Emitters drawn on the device (no reference counterpart): it replaces the
   host draws of simobj.sphere (simobj.py:119-130, util/random.py:39-51) and of
   the random phases (simobj.py:56-57).  Counter-based Philox4x32-10 with key
   `seed`; emitter i of `rng_stream` depends only on (seed, rng_stream, i).
     out:       n x 4 float64 (x, y, z, phi), written in place
     positions: 1 = uniform in the ball |r - centre| <= radius, 0 = keep x, y, z
     phases:    1 = phi ~ U[0, 2 pi), 0 = keep phi
     centre:    HOST array of 3 doubles or NULL (origin)                    */
IDISIM_API int idisim_generate_ball(uint64_t seed, uint64_t rng_stream, int64_t n, double radius, const double *centre,
                         int positions, int phases, double *out, void *stream);

/* Emitters in several balls of one radius (no reference counterpart; it moves
   the emitter draw of simobj.multisphere, simobj.py:230-262, onto the GPU; the
   caller draws the ball centres and per-ball counts).  Emitter i is drawn as by
   idisim_generate_ball with centre 0 and shifted by the centre of the ball b
   with offsets[b] <= i < offsets[b + 1].
     centres: nballs x 3 float64, DEVICE;  offsets: nballs + 1 int64, DEVICE,
     non-decreasing, offsets[0] = 0, offsets[nballs] = n                      */
IDISIM_API int idisim_generate_balls(uint64_t seed, uint64_t rng_stream, int64_t n, double radius, int64_t nballs,
                                     const double *centres, const int64_t *offsets, int positions, int phases,
                                     double *out, void *stream);

/* Emitters in a Gaussian cloud (no reference counterpart; it moves the draw of
   simobj.gauss / simobj.gnorm with rho = 2, simobj.py:154-184, onto the GPU):
   x = sigma[0] z0, y = sigma[1] z1, z = sigma[2] z2 with standard normal z
   (Box-Muller on Philox4x32-10), then p -> rot p if rot is not NULL.
     sigma: HOST array of 3 standard deviations;  rot: HOST 3 x 3 row-major or NULL */
IDISIM_API int idisim_generate_gauss(uint64_t seed, uint64_t rng_stream, int64_t n, const double *sigma,
                                     const double *rot, int positions, int phases, double *out, void *stream);

/* Emitters chosen without replacement from fixed sites (no reference counterpart; it
   moves the per-image draw of simobj.crystal, simobj.py:479-487, onto the GPU).  Site i
   gets the key -ln(u_i)/weights[i] (u_i uniform, Philox4x32-10 with key `seed`); the n
   smallest keys are chosen (Efraimidis-Spirakis: the distribution of n successive draws
   proportional to the remaining weights, as numpy's choice(replace=False, p=weights));
   n > nsites takes further rounds with fresh keys.  Chosen sites are written to
   out[:, 0:3] in key order, rotated by rot when given; phases = 1 also draws
   out[:, 3] ~ U[0, 2 pi), 0 keeps it.
     sites: nsites x 3 float64, weights: nsites float64 or NULL (uniform), both DEVICE;
     rot: HOST 3 x 3 row-major or NULL; out: n x 4 float64, device.
     workspace: device scratch of *workspace_bytes bytes; with workspace == NULL the
     required size is written to *workspace_bytes and nothing else is done.        */
IDISIM_API int idisim_generate_choice(uint64_t seed, uint64_t rng_stream, int64_t n, int64_t nsites,
                                      const double *sites, const double *weights, const double *rot, int phases,
                                      double *out, void *workspace, size_t *workspace_bytes, void *stream);

/* Charge sharing between neighbouring detector pixels (replaces chargesharing /
   __putcharges, common.py:26-67).  Every count of `counts` (n0 x n1 float64,
   device; truncated toward zero, non-positive values give no hits) is a hit at a
   uniform random position in its pixel whose charge, a Gaussian of width `sigma`
   pixels, is integrated over the (2w+1)^2 pixels around it, w = ceil(2 sigma);
   charge leaving the detector is lost.  gain_noise > 0 scales each hit by
   1 + gain_noise N(0,1).  sigma <= 0 (or NaN): out = counts (1 + gain_noise N(0,1))
   per pixel, or a copy of counts.  Random numbers: Philox4x32-10 with key `seed`
   (not numpy's stream).  out: n0 x n1 float64, device, must not overlap counts.
   sigma is limited to 32 pixels.                                            */
IDISIM_API int idisim_chargesharing(const double *counts, int64_t n0, int64_t n1, double sigma, double gain_noise,
                                    uint64_t seed, double *out, void *stream);

/* Time-integrated intensity of decaying emitters (replaces cutime.simulate,
   cutime.py:9-138 and simtime.cu:43-242, and the CPU time.simulate, time.py:47-98).
   Emitter j is excited at times[j] and radiates a_j exp(-(t - t_j)/tau) from its
   arrival time t_j at the pixel on; out[x, y] = integral of |E(t)|^2 dt.  Geometry,
   near/far field and "scale" come from the handle (offsets and rotation included).
   Arrival times are float32 on FP32-split handles (the reference's "mixed"
   precision) and float64 on FP64 handles; amplitudes and sums are FP64.
     emitters: n x 4 float64 (x, y, z, phi), times: n float64, both device;
     c: speed of light, tau: decay time (same units as times, > 0);
     out: n0 x n1 float64, device.  Synchronises nothing; uses handle scratch.   */
IDISIM_API int idisim_time_forward(idisim_handle *h, const double *emitters, const double *times, int64_t n,
                                   double c, double tau, double *out, void *stream);

/* One streaming step of the 2-d FFT correlator (replaces the division and
   running statistics of correlator2d.add, cusimple.py:80-107): per element
   c = corr / maskcorr where maskcorr > 0.5, else 0; count == 1 starts mean = c,
   m2 = 0, later counts update mean += (c - mean) / count and
   m2 += (c - mean_old)(c - mean_new).  All arrays: n float32, device.        */
IDISIM_API int idisim_corr_accumulate(const float *corr, const float *maskcorr, int64_t n, int64_t count,
                                      float *mean, float *m2, void *stream);

/* q-corrected autocorrelation of one nr x nc image (replaces the kernel and the
   assemble step of cucor.corrfunction, cucor.py:9-133).  zd: nr x nc float32
   z / |(r - xcenter, c - ycenter, z)|, img: nr x nc float32, both device.
   out: (2 maxdqz) x (2 qmax) x (2 qmax) float64, device, overwritten.
   Requires 1 <= qmax <= 1024, maxdqz >= 1.                                     */
IDISIM_API int idisim_qcorr(const float *zd, const float *img, int64_t nr, int64_t nc, double xcenter,
                            double ycenter, double z, int64_t qmax, int64_t maxdqz, double *out, void *stream);

/* Issue rates of the instructions that bound the FP32-split kernels, measured on
   `device`: v_fma_f32 and v_sin_f32 / v_cos_f32 lane operations per second, each
   alone.  The two do not overlap in issue, so work that takes f FMA-class and t
   transcendental lane operations per unit runs at most 1 / (f / fma + t / trans)
   units per second.  No counterpart in the reference; for callers that
   state a kernel's issue ceiling (the benchmark, autotuning across devices).
   Synchronous, a few milliseconds, on a stream of its own.                      */
IDISIM_API int idisim_issue_rates(int device, double *fma_per_s, double *trans_per_s);

#ifdef __cplusplus
}
#endif
#endif /* IDISIM_H */
