// C ABI of the MI355X direct-sum wavefield library (declared in include/idisim.h).
//
// Host-side responsibilities:
//   - geometry set-up in FP64 (replaces the closure of cpu.py:6-40 and the
//     per-call flag/compile step of cuda.py:53-86);
//   - the FP32-split error planner: for every tile variant, an a-priori bound
//     of the per-term phase error (FP32 rounding + nf linearisation + sin/cos)
//     from the detector geometry and the emitter extent; tiles whose bound
//     exceeds the tolerance are not used, and when none qualifies the handle
//     runs the FP64 kernel (or returns IDISIM_E_PRECISION);
//   - the autotuner: times candidate (tile, emitter split / blocks per CU)
//     plans with hipEvents on the caller's stream and caches the fastest per
//     (output kind, amplitudes, batched, log2 n);
//   - scratch, work-item tables, and hipGraph capture/replay of the
//     persistent batched launch.
#include "../../include/idisim.h"
#include "kernels.hpp"
#include "checks.hpp"
#include "planner.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

using namespace idisim;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define IDISIM_HIP(expr)                                                                      \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return fail(IDISIM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// A tuned choice is keyed by the call's shape and kernel family (FP32-split or
// FP64); it is reused only while its tile is still among the eligible ones (a
// larger emitter extent can push a tile's error bound past the tolerance).
struct TuneKey {
    int kind, has_amp, batched, split_path, nbucket;
    bool operator<(const TuneKey& o) const
    {
        return std::tie(kind, has_amp, batched, split_path, nbucket) <
               std::tie(o.kind, o.has_amp, o.batched, o.split_path, o.nbucket);
    }
};

static bool has_tile(const std::vector<int>& tiles, int v)
{
    return std::find(tiles.begin(), tiles.end(), v) != tiles.end();
}

struct Choice {
    int tile = 2;
    int split = 1;       // emitter splits (single shot) or blocks per CU (batched)
    double ms = 0.0;
};

}  // namespace

// The handle: the planner state (geometry, tolerance, tables; planner.hpp) and everything a
// call needs besides.
struct idisim_handle : idisim::Planner {
    int device = 0;
    int precision = IDISIM_PREC_FP32SPLIT;
    int num_cus = 256;
    idisim_geometry geom{};
    // options (tol_rev and summode are in the planner state)
    int autotune = 1, force_tile = -1, force_split = 0, use_graph = 1, fallback = 1;
    int time_path = 0;
    double rho = 0.0;   // emitter extent of the current call
    // device scratch
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    unsigned long long* d_ext = nullptr;
    // extent check: the FP32-split record stage sets the flag (pinned host memory, read by
    // the host without a synchronisation) when an emitter lies beyond the planner's extent
    unsigned int* h_viol = nullptr;   // host view
    unsigned int* d_viol = nullptr;   // device view of the same word
    int viol_seen = 0;                // the flag was found set (reported by idisim_get_plan)
    void* noalloc = nullptr;          // this device's address of the static no-alloc scratch
    // batched work items
    WorkItem* d_items = nullptr;
    size_t items_cap = 0;
    WorkItem* h_items = nullptr;  // pinned staging
    std::vector<int64_t> items_offsets;
    int items_tile = -1;
    int nitems = 0;
    int* d_counter = nullptr;
    // graph of the batched launch
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<uintptr_t> gkey;
    std::map<TuneKey, Choice> tuned;
    idisim_plan plan{};
    // set-up tables of the single-shot mirror-pair variants (geometry only), built on first use
    std::map<int, float*> pair_tabs;
    void* field_tmp = nullptr;  // idisim_loss_grad_l2's field when the caller passes none
};

namespace {

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

size_t out_elem_bytes(int precision, int kind)
{
    if (precision == IDISIM_PREC_FP64) return kind == IDISIM_OUT_INTENSITY ? sizeof(double) : 2 * sizeof(double);
    return kind == IDISIM_OUT_INTENSITY ? sizeof(float) : 2 * sizeof(float);
}

int ensure_scratch(idisim_handle* h, size_t bytes)
{
    if (bytes <= h->scratch_bytes) return IDISIM_OK;
    if (h->scratch) {
        (void)hipDeviceSynchronize();
        (void)hipFree(h->scratch);
        h->scratch = nullptr;
        h->scratch_bytes = 0;
    }
    if (hipMalloc(&h->scratch, bytes) != hipSuccess) {
        h->scratch = nullptr;
        return fail(IDISIM_E_NOMEM, "scratch allocation of " + std::to_string(bytes) + " bytes failed");
    }
    h->scratch_bytes = bytes;
    return IDISIM_OK;
}

// max_j |r_j| of the call's emitters, reduced into the device word `word` (the handle's own
// when NULL); synchronises the stream.
int measure_extent(idisim_handle* h, const double* emit, int64_t n, hipStream_t st, double* rho,
                   unsigned long long* word = nullptr)
{
    if (n <= 0) {
        *rho = 0.0;
        return IDISIM_OK;
    }
    if (word == nullptr) {
        if (!h->d_ext) IDISIM_HIP(hipMalloc(&h->d_ext, sizeof(unsigned long long)));
        word = h->d_ext;
    }
    IDISIM_HIP(launch_extent(emit, n, word, st));
    unsigned long long bits = 0;
    IDISIM_HIP(hipMemcpyAsync(&bits, word, sizeof(bits), hipMemcpyDeviceToHost, st));
    IDISIM_HIP(hipStreamSynchronize(st));
    double r2;
    std::memcpy(&r2, &bits, sizeof(r2));
    *rho = std::sqrt(r2);
    return IDISIM_OK;
}

int nbucket(int64_t n)
{
    int b = 0;
    while (n > 1) {
        n >>= 1;
        ++b;
    }
    return b;
}

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    Timer()
    {
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
    }
    ~Timer()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// Geometry of an FP32-split launch: the W fold decided for this tile and the call's extent,
// and the extent check of the record stage (|r_j| beyond the extent the bound assumed).
Geom split_geom(const idisim_handle* h, int tile, bool* w)
{
    Geom gl = h->g;
    *w = gl.nf || gl.secondorder;
    if (*w && fold_error(h, tile, h->rho) <= 0.1 * h->tol_rev) {
        gl.foldw = 1;
        *w = false;
    }
    // a relative margin far above the FP64 rounding of |r_j|^2 and of the caller's extent
    const double e = h->rho * (1.0 + 1e-9);
    gl.ext2 = e * e;
    gl.viol = h->d_viol;
    if (tile_ref_kind(tile) >= 4) {
        gl.refrho = (float)(h->rho * (1.0 + sigma_bound(h, tile, h->rho)));
        gl.refc = (float)ref_trunc_centre(h, tile, h->rho);
    }
    return gl;
}

// ------------------------------ single shot --------------------------------
struct FwdCall {
    const double* emit;
    const double* amp;
    int64_t n;
    void* out;
    int kind;
    hipStream_t st;
    bool split_path;  // FP32-split kernels (false: FP64 kernels)
    bool out_f32;     // FP64 kernels writing a complex64 / float32 output (fallback)
    // fixed device scratch of idisim_forward_noalloc (nullptr: the handle's own, grown on demand)
    void* fixed = nullptr;
    size_t fixed_cap = 0;
};

// Device scratch of `bytes` for one call: the fixed buffer of a no-alloc call, which is never
// grown, or the handle's buffer.
int call_scratch(idisim_handle* h, const FwdCall& c, size_t bytes, void** p)
{
    if (c.fixed != nullptr) {
        if (bytes > c.fixed_cap)
            return fail(IDISIM_E_INVALID, "call needs " + std::to_string(bytes) + " bytes of scratch, the static buffer holds " +
                                              std::to_string(c.fixed_cap));
        *p = c.fixed;
        return IDISIM_OK;
    }
    int rc = ensure_scratch(h, bytes);
    if (rc) return rc;
    *p = h->scratch;
    return IDISIM_OK;
}

// The set-up table of pair variant `tile`, built on the call's stream on first use (nullptr: the
// kernel forms the set-up in its prologue).  A no-alloc call builds none; tables above 64 MiB
// (detectors above ~2900^2 pixels) are not kept, as the prologue is then a small share of a launch.
int pair_table(idisim_handle* h, const FwdCall& c, int tile, const float** tab)
{
    *tab = nullptr;
    const size_t bytes = pair_table_bytes(tile, h->g);
    if (bytes == 0 || bytes > ((size_t)64 << 20)) return IDISIM_OK;
    auto it = h->pair_tabs.find(tile);
    if (it != h->pair_tabs.end()) {
        *tab = it->second;
        return IDISIM_OK;
    }
    if (c.fixed != nullptr) return IDISIM_OK;
    float* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return fail(IDISIM_E_NOMEM, "pair set-up table of " + std::to_string(bytes) + " bytes");
    hipError_t e = launch_pair_table(tile, h->g, p, c.st);
    if (e != hipSuccess) {
        (void)hipFree(p);
        return fail(IDISIM_E_HIP, std::string("launch_pair_table: ") + hipGetErrorString(e));
    }
    h->pair_tabs[tile] = p;
    *tab = p;
    return IDISIM_OK;
}

int run_forward(idisim_handle* h, const FwdCall& c, int tile, int splits)
{
    const Geom& g = h->g;
    const size_t npix = (size_t)g.n0 * g.n1;
    if (c.split_path) {
        bool w = false;
        const Geom gl = split_geom(h, tile, &w);
        const bool a = (c.amp != nullptr) || (g.nf && g.scale);
        float4* partial = nullptr;
        if (splits > 1) {
            void* p = nullptr;
            int rc = call_scratch(h, c, (size_t)splits * npix * sizeof(float4), &p);
            if (rc) return rc;
            partial = static_cast<float4*>(p);
        }
        if (tile == kTileNone) {
            return fail(IDISIM_E_UNSUPPORTED, "tile variant 4 has no kernel");
        } else if (tile_pair(tile)) {
            if (h->summode != 0) return fail(IDISIM_E_UNSUPPORTED, "the mirror-pair tiles have no per-term Kahan mode");
            if (w) return fail(IDISIM_E_UNSUPPORTED, "the mirror-pair tiles need the curvature folded into A");
            const float* tab = nullptr;
            if (int rc = pair_table(h, c, tile, &tab)) return rc;
            IDISIM_HIP(launch_fwd_pair(tile, a, gl, c.emit, c.amp, c.n, splits, tab, partial, c.out, c.kind, c.st));
        } else {
            IDISIM_HIP(launch_fwd_split(tile, w, a, h->summode, gl, c.emit, c.amp, c.n, splits, partial, c.out,
                                        c.kind, c.st));
        }
        if (splits > 1) IDISIM_HIP(launch_finalize_split(gl, partial, splits, c.out, c.kind, c.st));
        return IDISIM_OK;
    }
    // FP64 kernels; a complex64/float32 output (fallback of an FP32-split
    // handle) always goes through the finalize kernel, which converts.
    double4* partial = nullptr;
    if (splits > 1 || c.out_f32) {
        void* p = nullptr;
        int rc = call_scratch(h, c, (size_t)splits * npix * sizeof(double4), &p);
        if (rc) return rc;
        partial = static_cast<double4*>(p);
    }
    IDISIM_HIP(launch_fwd_fp64(tile, g, c.emit, c.amp, c.n, splits, partial, c.out, c.kind, c.st));
    if (partial != nullptr) {
        IDISIM_HIP(launch_finalize_fp64(g, partial, splits, c.out, c.kind | (c.out_f32 ? 2 : 0), c.st));
    }
    return IDISIM_OK;
}

// The mirror-pair forward of c into float4 split partials [splits][npix] only (no finalize):
// idisim_loss_grad_l2 sums them in its adjoint pixel stage.
int run_forward_partials(idisim_handle* h, const FwdCall& c, int tile, int splits, float4* partial)
{
    bool w = false;
    const Geom gl = split_geom(h, tile, &w);
    const bool a = (c.amp != nullptr) || (h->g.nf && h->g.scale);
    if (w) return fail(IDISIM_E_UNSUPPORTED, "the mirror-pair tiles need the curvature folded into A");
    const float* tab = nullptr;
    if (int rc = pair_table(h, c, tile, &tab)) return rc;
    IDISIM_HIP(launch_fwd_pair(tile, a, gl, c.emit, c.amp, c.n, splits, tab, partial, nullptr, c.kind, c.st));
    return IDISIM_OK;
}

std::vector<int> split_candidates(const idisim_handle* h, int64_t n, int tile)
{
    const int64_t maxs = std::max<int64_t>(1, (n + kSplitGrain - 1) / kSplitGrain);
    const int tiles = ((h->g.n0 + tile_rows(tile) - 1) / tile_rows(tile)) *
                      ((h->g.n1 + tile_cols(tile) - 1) / tile_cols(tile));
    std::vector<int> out;
    // powers of two, and 5 x 2^k for the kernels that hold five workgroups per CU (whole rounds)
    for (int s : {1, 2, 4, 5, 8, 10, 16, 20, 32, 40, 64, 80}) {
        if (s > maxs) break;
        out.push_back(s);
        if ((int64_t)tiles * s >= 20LL * h->num_cus) break;  // 20 workgroups per CU is plenty
    }
    return out;
}

int heuristic_split(const idisim_handle* h, int64_t n, int tile)
{
    const std::vector<int> c = split_candidates(h, n, tile);
    const int tiles = ((h->g.n0 + tile_rows(tile) - 1) / tile_rows(tile)) *
                      ((h->g.n1 + tile_cols(tile) - 1) / tile_cols(tile));
    for (int s : c)
        if ((int64_t)tiles * s >= 4LL * h->num_cus) return s;
    return c.back();
}

// The (tile, split) of a forward call: forced by the options, cached, autotuned (which runs
// candidate forwards into c.out), or the heuristic.
int plan_choose(idisim_handle* h, const FwdCall& c, const std::vector<int>& tiles, Choice& ch)
{
    const TuneKey key{c.kind, c.amp != nullptr, 0, c.split_path ? 1 : 0, nbucket(c.n)};
    auto it = h->tuned.find(key);
    const bool forced = h->force_tile >= 0 || h->force_split > 0;
    if (forced) {
        if (h->force_tile >= 0 && !has_tile(tiles, h->force_tile))
            return fail(IDISIM_E_UNSUPPORTED, "tile variant " + std::to_string(h->force_tile) + " has no FP64 kernel");
        ch.tile = h->force_tile >= 0 ? h->force_tile : tiles.front();
        ch.split = h->force_split > 0 ? h->force_split : heuristic_split(h, c.n, ch.tile);
    } else if (it != h->tuned.end() && has_tile(tiles, it->second.tile)) {
        ch = it->second;
    } else if (h->autotune && c.fixed == nullptr) {
        // time candidates on (at most) ~2e10 pairs of this call's work
        const size_t npix = (size_t)h->g.n0 * h->g.n1;
        int64_t nt = c.n;
        const double cap = c.split_path ? 2e10 : 2e9;
        if ((double)nt * (double)npix > cap) nt = std::max<int64_t>(kChunk, (int64_t)(cap / (double)npix));
        FwdCall ct = c;
        ct.n = nt;
        Timer tm;
        // Two passes, each candidate's fastest run over both; the second pass takes the candidates
        // within 25% of the best.  From idle, the GPU runs the first ~0.1 s of launches at a lower
        // clock (the same 4096-emitter forward measured 58 us cold and 51 us warm), so a single
        // pass favoured the later candidates.
        std::vector<std::pair<int, int>> cand;
        for (int v : tiles)
            for (int s : split_candidates(h, nt, v)) cand.emplace_back(v, s);
        std::vector<float> fastest(cand.size(), std::numeric_limits<float>::infinity());
        for (int pass = 0; pass < 2; ++pass) {
            const float cut = pass == 0 ? std::numeric_limits<float>::infinity()
                                        : 1.25f * *std::min_element(fastest.begin(), fastest.end());
            for (size_t i = 0; i < cand.size(); ++i) {
                if (fastest[i] > cut) continue;
                const int v = cand[i].first, s = cand[i].second;
                int rc = run_forward(h, ct, v, s);  // warm-up
                if (rc) return rc;
                // at least two runs (variants a few % apart), more for short launches, whose times
                // scatter by 10% from run to run (up to 8 runs: 0.25 ms)
                for (int rep = 0, reps = 2; rep < reps; ++rep) {
                    float m = 0.f;
                    IDISIM_HIP(hipEventRecord(tm.a, c.st));
                    rc = run_forward(h, ct, v, s);
                    if (rc) return rc;
                    IDISIM_HIP(hipEventRecord(tm.b, c.st));
                    IDISIM_HIP(hipEventSynchronize(tm.b));
                    IDISIM_HIP(hipEventElapsedTime(&m, tm.a, tm.b));
                    fastest[i] = std::min(fastest[i], m);
                    if (rep == 0) reps = std::max(2, std::min(8, (int)(0.25f / std::max(m, 1e-3f))));
                }
            }
        }
        float best = std::numeric_limits<float>::infinity();
        const bool log = std::getenv("IDISIM_TUNE_LOG") != nullptr;  // the candidates' times on stderr
        for (size_t i = 0; i < cand.size(); ++i) {
            if (log)
                std::fprintf(stderr, "idisim tune: n %lld tile %d split %d: %.4f ms\n", (long long)nt, cand[i].first,
                             cand[i].second, fastest[i]);
            if (fastest[i] < best) {
                best = fastest[i];
                ch.tile = cand[i].first;
                ch.split = cand[i].second;
                ch.ms = fastest[i];
            }
        }
        // the split that won on nt emitters, rescaled to n by workgroup count
        if (nt < c.n) {
            const std::vector<int> full = split_candidates(h, c.n, ch.tile);
            ch.split = std::min(ch.split, full.back());
        }
        h->tuned[key] = ch;
        // the set-up tables of the candidates that lost (another call may build one again)
        for (auto it2 = h->pair_tabs.begin(); it2 != h->pair_tabs.end();) {
            bool keep = it2->first == ch.tile;
            for (const auto& kv : h->tuned) keep = keep || kv.second.tile == it2->first;
            if (keep) {
                ++it2;
            } else {
                IDISIM_HIP(hipStreamSynchronize(c.st));
                (void)hipFree(it2->second);
                it2 = h->pair_tabs.erase(it2);
            }
        }
    } else {
        // heuristic: a mirror-pair quad per lane if allowed, enough splits for 4 WG/CU
        ch.tile = tiles.front();
        for (int v : {kTileRefL, kTileRef1, kTileRefC1, kTileRefSC, kTileRefS, kTileRef2, kTileRefC2, kTilePair1, kTileBlock, kTilePairC, 2, 1, 3, 0})
            if (has_tile(tiles, v)) {
                ch.tile = v;
                break;
            }
        ch.split = heuristic_split(h, c.n, ch.tile);
    }
    const int64_t maxs = std::max<int64_t>(1, (c.n + kSplitGrain - 1) / kSplitGrain);
    if (ch.split > maxs) ch.split = (int)maxs;
    if (c.fixed != nullptr && ch.split > 1) {
        // a no-alloc call: as many splits as the static buffer holds partials for
        const size_t per = (size_t)h->g.n0 * h->g.n1 * (c.split_path ? sizeof(float4) : sizeof(double4));
        ch.split = (int)std::max<size_t>(1, std::min<size_t>((size_t)ch.split, c.fixed_cap / per));
    }
    return IDISIM_OK;
}

// The plan a forward ran with, for idisim_get_plan.
void plan_record(idisim_handle* h, const FwdCall& c, const Choice& ch)
{
    if (c.split_path) h->plan.phase_bound = plan_bound(h, ch.tile, h->rho);  // the tile that ran
    h->plan.tile = ch.tile;
    h->plan.tile_rows = tile_rows(ch.tile);
    h->plan.tile_cols = tile_cols(ch.tile);
    h->plan.split = ch.split;
    h->plan.tuned_ms = ch.ms;
}

int plan_and_run(idisim_handle* h, FwdCall c, const std::vector<int>& tiles)
{
    Choice ch;
    if (int rc = plan_choose(h, c, tiles, ch)) return rc;
    if (int rc = run_forward(h, c, ch.tile, ch.split)) return rc;
    plan_record(h, c, ch);
    return IDISIM_OK;
}

// ------------------------------ batched ------------------------------------
int build_items(idisim_handle* h, const int64_t* offsets, int64_t nshots, int tile, hipStream_t st)
{
    const Geom& g = h->g;
    std::vector<int64_t> key(offsets, offsets + nshots + 1);
    if (tile == h->items_tile && key == h->items_offsets && h->d_items) return IDISIM_OK;
    const int tx = (g.n0 + tile_rows(tile) - 1) / tile_rows(tile);
    const int ty = (g.n1 + tile_cols(tile) - 1) / tile_cols(tile);
    const int64_t ntiles = (int64_t)tx * ty;
    const int64_t nitems = ntiles * nshots;
    if (nitems > std::numeric_limits<int>::max()) return fail(IDISIM_E_INVALID, "too many work items");
    std::vector<WorkItem> items;
    items.reserve((size_t)nitems);
    for (int64_t s = 0; s < nshots; ++s)
        for (int64_t t = 0; t < ntiles; ++t) items.push_back(WorkItem{(int32_t)s, (int32_t)t, offsets[s], offsets[s + 1]});
    // longest first: the queue tail is made of the shortest items
    std::stable_sort(items.begin(), items.end(),
                     [](const WorkItem& a, const WorkItem& b) { return (a.e1 - a.e0) > (b.e1 - b.e0); });
    IDISIM_HIP(hipStreamSynchronize(st));  // the staging buffer may still feed a previous copy
    if ((size_t)nitems > h->items_cap) {
        if (h->d_items) (void)hipFree(h->d_items);
        if (h->h_items) (void)hipHostFree(h->h_items);
        h->d_items = nullptr;
        h->h_items = nullptr;
        h->items_cap = 0;
        if (hipMalloc(&h->d_items, (size_t)nitems * sizeof(WorkItem)) != hipSuccess)
            return fail(IDISIM_E_NOMEM, "work item table allocation failed");
        if (hipHostMalloc(&h->h_items, (size_t)nitems * sizeof(WorkItem), hipHostMallocDefault) != hipSuccess)
            return fail(IDISIM_E_NOMEM, "pinned work item staging allocation failed");
        h->items_cap = (size_t)nitems;
    }
    std::memcpy(h->h_items, items.data(), items.size() * sizeof(WorkItem));
    IDISIM_HIP(hipMemcpyAsync(h->d_items, h->h_items, items.size() * sizeof(WorkItem), hipMemcpyHostToDevice, st));
    if (!h->d_counter) IDISIM_HIP(hipMalloc(&h->d_counter, sizeof(int)));
    h->items_offsets = std::move(key);
    h->items_tile = tile;
    h->nitems = (int)nitems;
    return IDISIM_OK;
}

struct BatchCall {
    const double* emit;
    const double* amp;
    void* out;
    int kind;
    hipStream_t st;
    bool split_path;
};

int launch_batched(idisim_handle* h, const BatchCall& c, int tile, int blocks_per_cu, hipStream_t st)
{
    const Geom& g = h->g;
    const int nblocks = std::max(1, blocks_per_cu * h->num_cus);
    if (c.split_path) {
        bool w = false;
        const Geom gl = split_geom(h, tile, &w);
        const bool a = (c.amp != nullptr) || (g.nf && g.scale);
        if (tile == kTileNone) {
            return fail(IDISIM_E_UNSUPPORTED, "tile variant 4 has no kernel");
        } else if (tile_pair(tile)) {
            if (h->summode != 0) return fail(IDISIM_E_UNSUPPORTED, "the mirror-pair tiles have no per-term Kahan mode");
            if (w) return fail(IDISIM_E_UNSUPPORTED, "the mirror-pair tiles need the curvature folded into A");
            IDISIM_HIP(launch_batched_pair(tile, a, gl, c.emit, c.amp, h->d_items, h->nitems, h->d_counter, nblocks,
                                           c.out, c.kind, st));
        } else {
            IDISIM_HIP(launch_batched_split(tile, w, a, h->summode, gl, c.emit, c.amp, h->d_items, h->nitems,
                                            h->d_counter, nblocks, c.out, c.kind, st));
        }
    } else {
        IDISIM_HIP(launch_batched_fp64(tile, g, c.emit, c.amp, h->d_items, h->nitems, h->d_counter, nblocks, c.out,
                                       c.kind, st));
    }
    return IDISIM_OK;
}

int occupancy(idisim_handle* h, const BatchCall& c, int tile)
{
    const Geom& g = h->g;
    if (c.split_path) {
        const bool w = (g.nf || g.secondorder) && fold_error(h, tile, h->rho) > 0.1 * h->tol_rev;
        if (tile_pair(tile)) return occupancy_batched_pair(tile, (c.amp != nullptr) || (g.nf && g.scale));
        return occupancy_batched_split(tile, w, (c.amp != nullptr) || (g.nf && g.scale), h->summode);
    }
    return occupancy_batched_fp64(tile);
}

int run_batched(idisim_handle* h, const BatchCall& c, int tile, int bpc)
{
    if (!h->use_graph) {
        h->plan.graph = 0;
        return launch_batched(h, c, tile, bpc, c.st);
    }
    // the W-fold decision and the extent check depend on the emitter extent of this call: a
    // graph captured with another extent must not be replayed
    const bool folded = c.split_path && (h->g.nf || h->g.secondorder) &&
                        fold_error(h, tile, h->rho) <= 0.1 * h->tol_rev;
    uint64_t rho_bits = 0;
    std::memcpy(&rho_bits, &h->rho, sizeof(rho_bits));
    std::vector<uintptr_t> key = {(uintptr_t)c.emit,    (uintptr_t)c.amp,       (uintptr_t)c.out,
                                  (uintptr_t)c.kind,    (uintptr_t)h->d_items,  (uintptr_t)h->nitems,
                                  (uintptr_t)tile,      (uintptr_t)bpc,         (uintptr_t)h->summode,
                                  (uintptr_t)c.split_path, (uintptr_t)folded, (uintptr_t)rho_bits};
    if (key != h->gkey || !h->gexec) {
        if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
        if (h->graph) (void)hipGraphDestroy(h->graph);
        h->gexec = nullptr;
        h->graph = nullptr;
        h->gkey.clear();
        if (!h->cap_stream) IDISIM_HIP(hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking));
        IDISIM_HIP(hipStreamBeginCapture(h->cap_stream, hipStreamCaptureModeThreadLocal));
        int rc = launch_batched(h, c, tile, bpc, h->cap_stream);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(h->cap_stream, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return fail(IDISIM_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        h->graph = graph;
        IDISIM_HIP(hipGraphInstantiate(&h->gexec, graph, nullptr, nullptr, 0));
        h->gkey = key;
    }
    IDISIM_HIP(hipGraphLaunch(h->gexec, c.st));
    h->plan.graph = 1;
    return IDISIM_OK;
}

int fill_zero(void* out, size_t bytes, hipStream_t st)
{
    IDISIM_HIP(launch_zero(out, bytes, st));
    return IDISIM_OK;
}

// An earlier FP32-split call of this handle met an emitter beyond the extent its error bound
// assumed: its kernels set the flag.  Reading it needs no synchronisation, so a kernel that
// is still running is seen by a later call.  The flag is cleared once reported.
int check_extent_flag(idisim_handle* h, const char* who)
{
    h->viol_seen = 0;
    if (h->h_viol == nullptr || *reinterpret_cast<volatile unsigned int*>(h->h_viol) == 0u) return IDISIM_OK;
    *reinterpret_cast<volatile unsigned int*>(h->h_viol) = 0u;
    h->viol_seen = 1;
    return fail(IDISIM_E_PRECISION, std::string(who) +
                                        ": an earlier call of this handle was given an extent below max |r_j| of its "
                                        "emitters, so the FP32-split error bound of its result did not hold; pass "
                                        "extent <= 0 to have it measured");
}

}  // namespace

// Static device scratch of idisim_forward_noalloc: the partial sums of the emitter splits
// (float4 or double4 per pixel and split).  Each device holds its own copy; host code uses
// only the address that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[IDISIM_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (current), resolved once per device.
int noalloc_scratch(int device, void** p)
{
    if (device < 0 || device >= kMaxDevices) return fail(IDISIM_E_INVALID, "device index beyond the static scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device] == nullptr) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || ptr == nullptr)
            return fail(IDISIM_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = ptr;
    }
    *p = g_scratch_ptr[device];
    return IDISIM_OK;
}

}  // namespace

// ============================== C ABI =======================================
extern "C" {

int idisim_version(void) { return IDISIM_VERSION; }

int idisim_issue_rates(int device, double* fma_per_s, double* trans_per_s)
{
    g_err.clear();
    if (!fma_per_s || !trans_per_s) return fail(IDISIM_E_INVALID, "idisim_issue_rates: NULL output");
    int ndev = 0;
    IDISIM_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(IDISIM_E_INVALID, "idisim_issue_rates: no such device");
    DeviceGuard guard(device);
    int cus = 0;
    IDISIM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    const int blocks = 16 * std::max(cus, 1);  // two rounds of 32 waves per CU
    struct Res {
        float* out = nullptr;
        hipStream_t st = nullptr;
        hipEvent_t a = nullptr, b = nullptr;
        ~Res()
        {
            if (st) (void)hipStreamSynchronize(st);
            if (out) (void)hipFree(out);
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
            if (st) (void)hipStreamDestroy(st);
        }
    } r;
    IDISIM_HIP(hipMalloc(&r.out, (size_t)blocks * 256 * sizeof(float)));
    IDISIM_HIP(hipStreamCreate(&r.st));
    IDISIM_HIP(hipEventCreate(&r.a));
    IDISIM_HIP(hipEventCreate(&r.b));
    double rate[2] = {0.0, 0.0};
    for (int kind = 0; kind < 2; ++kind) {
        const int iters = kind == 0 ? 1024 : 512;  // ~1 ms per launch on 256 CUs
        const double ops = (double)blocks * 256.0 * iters * (kind == 0 ? 64.0 : 32.0);
        IDISIM_HIP(launch_issue_probe(kind, blocks, iters, r.out, r.st));  // warm-up
        for (int rep = 0; rep < 3; ++rep) {
            IDISIM_HIP(hipEventRecord(r.a, r.st));
            IDISIM_HIP(launch_issue_probe(kind, blocks, iters, r.out, r.st));
            IDISIM_HIP(hipEventRecord(r.b, r.st));
            IDISIM_HIP(hipEventSynchronize(r.b));
            float ms = 0.f;
            IDISIM_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            if (ms > 0.f) rate[kind] = std::max(rate[kind], ops / (ms * 1e-3));
        }
    }
    *fma_per_s = rate[0];
    *trans_per_s = rate[1];
    return IDISIM_OK;
}

const char* idisim_last_error(void) { return g_err.c_str(); }

int idisim_create(idisim_handle** out, const idisim_geometry* geom, int precision, int device)
{
    g_err.clear();
    if (!out) return fail(IDISIM_E_INVALID, "idisim_create: NULL argument");
    *out = nullptr;
    std::string msg;
    if (int rc = check_create(geom, precision, &msg)) return fail(rc, msg);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(IDISIM_E_HIP, "idisim_create: no HIP device available");
    if (device < 0 || device >= ndev) return fail(IDISIM_E_INVALID, "idisim_create: invalid device");
    DeviceGuard guard(device);
    idisim_handle* h = new idisim_handle();
    h->device = device;
    h->precision = precision;
    h->geom = *geom;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        h->num_cus = prop.multiProcessorCount;
    h->g = make_geom(*geom);
    planner_tables(h);
    h->plan.precision = precision;
    if (hipHostMalloc(reinterpret_cast<void**>(&h->h_viol), sizeof(unsigned int),
                      hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
        h->h_viol = nullptr;
        delete h;
        return fail(IDISIM_E_NOMEM, "idisim_create: pinned flag allocation failed");
    }
    *h->h_viol = 0u;
    if (hipHostGetDevicePointer(reinterpret_cast<void**>(&h->d_viol), h->h_viol, 0) != hipSuccess) {
        (void)hipHostFree(h->h_viol);
        delete h;
        return fail(IDISIM_E_HIP, "idisim_create: hipHostGetDevicePointer failed");
    }
    // the static scratch of idisim_forward_noalloc, resolved here once per device
    int rc = noalloc_scratch(device, &h->noalloc);
    if (rc) {
        (void)hipHostFree(h->h_viol);
        delete h;
        return rc;
    }
    *out = h;
    return IDISIM_OK;
}

int idisim_destroy(idisim_handle* h)
{
    if (!h) return IDISIM_OK;
    DeviceGuard guard(h->device);
    (void)hipDeviceSynchronize();
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    if (h->cap_stream) (void)hipStreamDestroy(h->cap_stream);
    if (h->scratch) (void)hipFree(h->scratch);
    if (h->d_ext) (void)hipFree(h->d_ext);
    if (h->d_items) (void)hipFree(h->d_items);
    if (h->h_items) (void)hipHostFree(h->h_items);
    if (h->d_counter) (void)hipFree(h->d_counter);
    if (h->h_viol) (void)hipHostFree(h->h_viol);
    for (auto& kv : h->pair_tabs)
        if (kv.second) (void)hipFree(kv.second);
    if (h->field_tmp) (void)hipFree(h->field_tmp);
    delete h;
    return IDISIM_OK;
}

int idisim_set_option(idisim_handle* h, int option, int64_t value)
{
    if (!h) return fail(IDISIM_E_INVALID, "idisim_set_option: NULL handle");
    std::string msg;
    if (int rc = check_option(option, value, kNumTiles, &msg)) return fail(rc, msg);
    switch (option) {
        case IDISIM_OPT_AUTOTUNE: h->autotune = value ? 1 : 0; break;
        case IDISIM_OPT_TILE: h->force_tile = (int)value; break;
        case IDISIM_OPT_SPLIT: h->force_split = (int)value; break;
        case IDISIM_OPT_SUMMODE: h->summode = (int)value; break;
        case IDISIM_OPT_GRAPH: h->use_graph = value ? 1 : 0; break;
        case IDISIM_OPT_TOL_PPB: h->tol_rev = (double)value * 1e-9; break;
        case IDISIM_OPT_FALLBACK: h->fallback = value ? 1 : 0; break;
        case IDISIM_OPT_TIME_PATH: h->time_path = (int)value; break;
        default: break;
    }
    h->tuned.clear();
    return IDISIM_OK;
}

int idisim_get_plan(const idisim_handle* h, idisim_plan* out)
{
    if (!h || !out) return fail(IDISIM_E_INVALID, "idisim_get_plan: NULL argument");
    *out = h->plan;
    out->extent_exceeded =
        (h->viol_seen || (h->h_viol != nullptr && *reinterpret_cast<volatile unsigned int*>(h->h_viol) != 0u)) ? 1 : 0;
    return IDISIM_OK;
}

int idisim_extent(const double* emitters, int64_t n, double* extent_out, void* stream)
{
    if (!extent_out || (n > 0 && !emitters) || n < 0) return fail(IDISIM_E_INVALID, "idisim_extent: bad argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (n == 0) {
        *extent_out = 0.0;
        return IDISIM_OK;
    }
    unsigned long long* d = nullptr;
    IDISIM_HIP(hipMalloc(&d, sizeof(unsigned long long)));
    hipError_t e = launch_extent(emitters, n, d, st);
    unsigned long long bits = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&bits, d, sizeof(bits), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(IDISIM_E_HIP, std::string("idisim_extent: ") + hipGetErrorString(e));
    double r2;
    std::memcpy(&r2, &bits, sizeof(r2));
    *extent_out = std::sqrt(r2);
    return IDISIM_OK;
}

int idisim_generate_ball(uint64_t seed, uint64_t rng_stream, int64_t n, double radius, const double* centre,
                         int positions, int phases, double* out, void* stream)
{
    g_err.clear();
    std::string msg;
    if (int rc = check_generate("idisim_generate_ball", n, out, radius, &msg)) return fail(rc, msg);
    if (centre && !(std::isfinite(centre[0]) && std::isfinite(centre[1]) && std::isfinite(centre[2])))
        return fail(IDISIM_E_INVALID, "idisim_generate_ball: centre must be finite");
    IDISIM_HIP(launch_gen_ball(seed, rng_stream, n, radius, centre, positions ? 1 : 0, phases ? 1 : 0, out,
                               static_cast<hipStream_t>(stream)));
    return IDISIM_OK;
}

int idisim_generate_balls(uint64_t seed, uint64_t rng_stream, int64_t n, double radius, int64_t nballs,
                          const double* centres, const int64_t* offsets, int positions, int phases, double* out,
                          void* stream)
{
    g_err.clear();
    std::string msg;
    if (int rc = check_generate("idisim_generate_balls", n, out, radius, &msg)) return fail(rc, msg);
    if (positions && n > 0 && (nballs < 1 || !centres || !offsets))
        return fail(IDISIM_E_INVALID, "idisim_generate_balls: positions need nballs >= 1, centres and offsets");
    IDISIM_HIP(launch_gen_balls(seed, rng_stream, n, radius, nballs, centres, offsets, positions ? 1 : 0,
                                phases ? 1 : 0, out, static_cast<hipStream_t>(stream)));
    return IDISIM_OK;
}

int idisim_generate_gauss(uint64_t seed, uint64_t rng_stream, int64_t n, const double* sigma, const double* rot,
                          int positions, int phases, double* out, void* stream)
{
    g_err.clear();
    if (n < 0 || (n > 0 && !out) || (positions && !sigma))
        return fail(IDISIM_E_INVALID, "idisim_generate_gauss: bad arguments");
    if (positions)
        for (int k = 0; k < 3; ++k)
            if (!(sigma[k] >= 0.0) || !std::isfinite(sigma[k]))
                return fail(IDISIM_E_INVALID, "idisim_generate_gauss: sigma must be finite and >= 0");
    if (rot)
        for (int k = 0; k < 9; ++k)
            if (!std::isfinite(rot[k])) return fail(IDISIM_E_INVALID, "idisim_generate_gauss: rot must be finite");
    const double zero[3] = {0.0, 0.0, 0.0};
    IDISIM_HIP(launch_gen_gauss(seed, rng_stream, n, sigma ? sigma : zero, rot, positions ? 1 : 0, phases ? 1 : 0,
                                out, static_cast<hipStream_t>(stream)));
    return IDISIM_OK;
}

int idisim_time_forward(idisim_handle* h, const double* emitters, const double* times, int64_t n, double c,
                        double tau, double* out, void* stream)
{
    g_err.clear();
    std::string msg;
    if (int rc = check_time_forward(h, emitters, times, n, c, tau, out, &msg)) return fail(rc, msg);
    DeviceGuard guard(h->device);
    const TimeLayout L = time_layout(h->g, n, h->precision == IDISIM_PREC_FP64, (size_t)1 << 30);
    if (n > 0 && !time_fused_ok(n, L.fp64, h->time_path)) {
        int rc = ensure_scratch(h, L.bytes);
        if (rc) return rc;
    }
    IDISIM_HIP(launch_time(h->g, L, emitters, times, n, c, tau, h->time_path, h->scratch, out,
                           static_cast<hipStream_t>(stream)));
    return IDISIM_OK;
}

int idisim_corr_accumulate(const float* corr, const float* maskcorr, int64_t n, int64_t count, float* mean,
                           float* m2, void* stream)
{
    g_err.clear();
    if (n < 0 || count < 1 || (n > 0 && (!corr || !maskcorr || !mean || !m2)))
        return fail(IDISIM_E_INVALID, "idisim_corr_accumulate: bad arguments");
    IDISIM_HIP(launch_corr_accumulate(corr, maskcorr, n, count, mean, m2, static_cast<hipStream_t>(stream)));
    return IDISIM_OK;
}

int idisim_qcorr(const float* zd, const float* img, int64_t nr, int64_t nc, double xcenter, double ycenter, double z,
                 int64_t qmax, int64_t maxdqz, double* out, void* stream)
{
    g_err.clear();
    std::string msg;
    if (int rc = check_qcorr(zd, img, nr, nc, xcenter, ycenter, z, qmax, maxdqz, out, &msg)) return fail(rc, msg);
    IDISIM_HIP(launch_qcorr(zd, img, (int)nr, (int)nc, (float)xcenter, (float)ycenter, (float)z, (int)qmax,
                            (int)maxdqz, out, static_cast<hipStream_t>(stream)));
    return IDISIM_OK;
}

int idisim_generate_choice(uint64_t seed, uint64_t rng_stream, int64_t n, int64_t nsites, const double* sites,
                           const double* weights, const double* rot, int phases, double* out, void* workspace,
                           size_t* workspace_bytes, void* stream)
{
    g_err.clear();
    if (!workspace_bytes) return fail(IDISIM_E_INVALID, "idisim_generate_choice: NULL workspace_bytes");
    if (n < 0 || nsites < 1 || nsites >= ((int64_t)1 << 31))
        return fail(IDISIM_E_INVALID, "idisim_generate_choice: need n >= 0 and 1 <= nsites < 2^31");
    const size_t need = choice_workspace_bytes(nsites);
    if (!workspace) {
        *workspace_bytes = need;
        return IDISIM_OK;
    }
    if (*workspace_bytes < need)
        return fail(IDISIM_E_INVALID, "idisim_generate_choice: workspace smaller than " + std::to_string(need) +
                                          " bytes");
    if (n > 0 && (!sites || !out)) return fail(IDISIM_E_INVALID, "idisim_generate_choice: NULL buffer");
    if (rot)
        for (int k = 0; k < 9; ++k)
            if (!std::isfinite(rot[k])) return fail(IDISIM_E_INVALID, "idisim_generate_choice: rot must be finite");
    if (n == 0) return IDISIM_OK;
    IDISIM_HIP(launch_choice(seed, rng_stream, n, nsites, sites, weights, rot, phases ? 1 : 0, out, workspace,
                             static_cast<hipStream_t>(stream)));
    return IDISIM_OK;
}

int idisim_chargesharing(const double* counts, int64_t n0, int64_t n1, double sigma, double gain_noise,
                         uint64_t seed, double* out, void* stream)
{
    g_err.clear();
    std::string msg;
    if (int rc = check_chargesharing(counts, n0, n1, sigma, gain_noise, out, &msg)) return fail(rc, msg);
    IDISIM_HIP(launch_chargesharing(counts, (int)n0, (int)n1, sigma, gain_noise, seed, out,
                                    static_cast<hipStream_t>(stream)));
    return IDISIM_OK;
}

// Decide FP32-split vs FP64 for a call; sets the plan's precision fields.
static int choose_path(idisim_handle* h, double rho, std::vector<int>* tiles, bool* split_path)
{
    h->rho = rho;
    h->plan.fallback = 0;
    h->plan.summode = h->summode;
    h->plan.graph = 0;
    if (h->precision == IDISIM_PREC_FP64) {
        *split_path = false;
        *tiles = {0, 1, 2, 3};
        h->plan.precision = IDISIM_PREC_FP64;
        h->plan.phase_bound = 0.0;
        return IDISIM_OK;
    }
    double best = 0.0;
    std::vector<int> ok = eligible_tiles(h, rho, &best);
    if (h->force_tile >= 0) {
        h->plan.phase_bound = plan_bound(h, h->force_tile, rho);
        *tiles = {h->force_tile};
        *split_path = true;
        h->plan.precision = IDISIM_PREC_FP32SPLIT;
        return IDISIM_OK;
    }
    if (ok.empty()) {
        if (!h->fallback)
            return fail(IDISIM_E_PRECISION, "FP32-split phase error bound " + std::to_string(best) +
                                                " rev exceeds the tolerance " + std::to_string(h->tol_rev) +
                                                " rev for this geometry and emitter extent");
        *split_path = false;
        *tiles = {0, 1, 2, 3};
        h->plan.fallback = 1;
        h->plan.precision = IDISIM_PREC_FP64;
        h->plan.phase_bound = 0.0;
        return IDISIM_OK;
    }
    double b = std::numeric_limits<double>::infinity();
    for (int v : ok) b = std::min(b, plan_bound(h, v, rho));
    h->plan.phase_bound = b;
    *tiles = ok;
    *split_path = true;
    h->plan.precision = IDISIM_PREC_FP32SPLIT;
    return IDISIM_OK;
}

// idisim_forward (noalloc false) and idisim_forward_noalloc (true: the partial sums and the
// extent word live in the static scratch, the plan is the tuned or heuristic one without a
// timing pass, and the emitter split is capped by the scratch size).
static int forward_call(idisim_handle* h, const double* emitters, const double* amplitudes, int64_t n, double extent,
                        void* out, int out_kind, hipStream_t st, bool noalloc)
{
    const char* who = noalloc ? "idisim_forward_noalloc" : "idisim_forward";
    DeviceGuard guard(h->device);
    int rc = check_extent_flag(h, who);
    if (rc) return rc;
    const size_t npix = (size_t)h->g.n0 * h->g.n1;
    if (n == 0) return fill_zero(out, npix * out_elem_bytes(h->precision, out_kind), st);
    double rho = extent;
    if (!(rho > 0.0)) {
        rc = measure_extent(h, emitters, n, st, &rho,
                            noalloc ? static_cast<unsigned long long*>(h->noalloc) : nullptr);
        if (rc) return rc;
    }
    std::vector<int> tiles;
    bool split_path = true;
    rc = choose_path(h, rho, &tiles, &split_path);
    if (rc) return rc;
    FwdCall c{emitters, amplitudes, n, out, out_kind, st, split_path,
              h->precision == IDISIM_PREC_FP32SPLIT && !split_path};
    if (noalloc) {
        c.fixed = h->noalloc;
        c.fixed_cap = IDISIM_NOALLOC_SCRATCH_BYTES;
    }
    return plan_and_run(h, c, tiles);
}

int idisim_forward(idisim_handle* h, const double* emitters, const double* amplitudes, int64_t n, double extent,
                   void* out, int out_kind, void* stream)
{
    g_err.clear();
    std::string msg;
    if (int rc = check_forward("idisim_forward", h, emitters, n, out, out_kind, extent, &msg)) return fail(rc, msg);
    return forward_call(h, emitters, amplitudes, n, extent, out, out_kind, static_cast<hipStream_t>(stream), false);
}

int idisim_forward_noalloc(idisim_handle* h, const double* emitters, const double* amplitudes, int64_t n,
                           double extent, void* out, int out_kind, void* stream)
{
    g_err.clear();
    std::string msg;
    if (int rc = check_forward("idisim_forward_noalloc", h, emitters, n, out, out_kind, extent, &msg))
        return fail(rc, msg);
    if (h->noalloc == nullptr) return fail(IDISIM_E_HIP, "idisim_forward_noalloc: no static scratch on this device");
    return forward_call(h, emitters, amplitudes, n, extent, out, out_kind, static_cast<hipStream_t>(stream), true);
}

int idisim_forward_batched(idisim_handle* h, const double* emitters, const double* amplitudes,
                           const int64_t* offsets, int64_t nshots, double extent, void* out, int out_kind,
                           void* stream)
{
    g_err.clear();
    std::string msg;
    if (int rc = check_batched(h, emitters, offsets, nshots, out, out_kind, extent, &msg)) return fail(rc, msg);
    const int64_t total = offsets[nshots];
    DeviceGuard guard(h->device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (int rc0 = check_extent_flag(h, "idisim_forward_batched")) return rc0;
    double rho = extent;
    if (!(rho > 0.0)) {
        int rc = measure_extent(h, emitters, total, st, &rho);
        if (rc) return rc;
    }
    std::vector<int> tiles;
    bool split_path = true;
    int rc = choose_path(h, rho, &tiles, &split_path);
    if (rc) return rc;
    if (h->precision == IDISIM_PREC_FP32SPLIT && !split_path) {
        // No tile meets the bound: the FP64 kernel shot by shot, writing the
        // complex64 / float32 output of this handle (no graph).
        const size_t shot_bytes =
            (size_t)h->g.n0 * h->g.n1 * out_elem_bytes(IDISIM_PREC_FP32SPLIT, out_kind);
        for (int64_t s = 0; s < nshots; ++s) {
            const int64_t ns = offsets[s + 1] - offsets[s];
            void* o = static_cast<char*>(out) + (size_t)s * shot_bytes;
            if (ns == 0) {
                rc = fill_zero(o, shot_bytes, st);
            } else {
                FwdCall fc{emitters + 4 * offsets[s], amplitudes ? amplitudes + offsets[s] : nullptr, ns, o,
                           out_kind, st, false, true};
                rc = run_forward(h, fc, 1, heuristic_split(h, ns, 1));
            }
            if (rc) return rc;
        }
        h->plan.tile = 1;
        h->plan.tile_rows = tile_rows(1);
        h->plan.tile_cols = tile_cols(1);
        h->plan.split = 1;
        h->plan.tuned_ms = 0.0;
        return IDISIM_OK;
    }
    BatchCall c{emitters, amplitudes, out, out_kind, st, split_path};

    const TuneKey key{out_kind, amplitudes != nullptr, 1, split_path ? 1 : 0,
                      nbucket(std::max<int64_t>(1, total / nshots))};
    Choice ch;
    auto it = h->tuned.find(key);
    if (h->force_tile >= 0 || h->force_split > 0) {
        if (h->force_tile >= 0 && !has_tile(tiles, h->force_tile))
            return fail(IDISIM_E_UNSUPPORTED, "tile variant " + std::to_string(h->force_tile) + " has no FP64 kernel");
        ch.tile = h->force_tile >= 0 ? h->force_tile : tiles.front();
        ch.split = h->force_split > 0 ? h->force_split : occupancy(h, c, ch.tile);
    } else if (it != h->tuned.end() && has_tile(tiles, it->second.tile)) {
        ch = it->second;
    } else if (h->autotune) {
        // time the candidates on a prefix of the batch of at most ~2e11 pairs (2e10 for the
        // FP64 kernels); a shot that crosses the cap is cut
        const double npix = (double)h->g.n0 * (double)h->g.n1;
        const double cap = split_path ? 2e11 : 2e10;
        std::vector<int64_t> toff(1, 0);
        for (int64_t s = 0; s < nshots; ++s) {
            int64_t end = offsets[s + 1];
            const double before = (double)offsets[s] * npix;
            if (before + (double)(end - offsets[s]) * npix > cap) {
                const int64_t room = std::max<int64_t>(kChunk, (int64_t)((cap - before) / npix));
                toff.push_back(std::min(end, offsets[s] + room));
                break;
            }
            toff.push_back(end);
        }
        const int64_t tshots = (int64_t)toff.size() - 1;
        Timer tm;
        double best = std::numeric_limits<double>::infinity();
        for (int v : tiles) {
            rc = build_items(h, toff.data(), tshots, v, st);
            if (rc) return rc;
            const int occ = occupancy(h, c, v);
            std::vector<int> bpcs = {occ};
            if (occ > 2) bpcs.push_back(occ / 2);
            for (int bpc : bpcs) {
                rc = launch_batched(h, c, v, bpc, st);  // warm-up
                if (rc) return rc;
                float ms = 0.f;
                IDISIM_HIP(hipEventRecord(tm.a, st));
                rc = launch_batched(h, c, v, bpc, st);
                if (rc) return rc;
                IDISIM_HIP(hipEventRecord(tm.b, st));
                IDISIM_HIP(hipEventSynchronize(tm.b));
                IDISIM_HIP(hipEventElapsedTime(&ms, tm.a, tm.b));
                if (ms < best) {
                    best = ms;
                    ch.tile = v;
                    ch.split = bpc;
                    ch.ms = ms;
                }
            }
        }
        h->tuned[key] = ch;
    } else {
        ch.tile = tiles.front();
        for (int v : {kTileRefL, kTileRef1, kTileRefC1, kTileRefSC, kTileRefS, kTileRef2, kTileRefC2, kTilePair1, kTileBlock, kTilePairC, 2, 1, 3, 0})
            if (has_tile(tiles, v)) {
                ch.tile = v;
                break;
            }
        ch.split = occupancy(h, c, ch.tile);
    }
    rc = build_items(h, offsets, nshots, ch.tile, st);
    if (rc) return rc;
    rc = run_batched(h, c, ch.tile, ch.split);
    if (rc) return rc;
    if (split_path) h->plan.phase_bound = plan_bound(h, ch.tile, h->rho);
    h->plan.tile = ch.tile;
    h->plan.tile_rows = tile_rows(ch.tile);
    h->plan.tile_cols = tile_cols(ch.tile);
    h->plan.split = ch.split;
    h->plan.tuned_ms = ch.ms;
    return IDISIM_OK;
}

// The forward that idisim_loss_grad_l2 runs inside its adjoint call: its planned mirror-pair
// tile and split; the field is summed from its partials (field may then be NULL: not stored).
struct FusedForward {
    FwdCall c;
    Choice ch;
};

// idisim_adjoint (target NULL), idisim_adjoint_l2 and (fz) idisim_loss_grad_l2
static int adjoint_call(const char* who, idisim_handle* h, const double* emitters, const double* amplitudes, int64_t n,
                        double extent, const void* field, const float* dLdI, const float* target, double* loss,
                        float* grad_pos, float* grad_phase, float* grad_amp, void* stream,
                        const FusedForward* fz = nullptr)
{
    g_err.clear();
    std::string msg;
    if (int rc = check_adjoint(who, h, emitters, n, fz ? static_cast<const void*>(target) : field,
                               target ? target : dLdI, grad_pos, grad_phase, extent, &msg))
        return fail(rc, msg);
    if (h->precision != IDISIM_PREC_FP32SPLIT)
        return fail(IDISIM_E_UNSUPPORTED, std::string(who) + ": only FP32-split handles have an adjoint");
    DeviceGuard guard(h->device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (n == 0) {
        if (target != nullptr && loss != nullptr) {
            // no emitters: E = 0 and the loss is sum T^2; no gradients to write
            return fail(IDISIM_E_INVALID, std::string(who) + ": the loss needs at least one emitter");
        }
        return IDISIM_OK;
    }
    if (int rc0 = check_extent_flag(h, who)) return rc0;
    double rho = extent;
    if (!(rho > 0.0)) {
        int rc = measure_extent(h, emitters, n, st, &rho);
        if (rc) return rc;
    }
    // the adjoint streams 16 x 16 pixel tiles (the planner's tile variant 0 has that shape); it
    // adds the per-pixel phase psi_p (FP32, |psi| <= 1/2) to the reduced phase: one more rounding
    // at |tau + psi| <= 1 rev
    constexpr int adj_tile = 0;
    const double bound = plan_bound(h, adj_tile, rho) + 0x1p-23;
    if (bound > h->tol_rev)
        return fail(IDISIM_E_PRECISION, std::string(who) + ": FP32-split error bound " + std::to_string(bound) +
                                            " rev exceeds the tolerance for this geometry");
    h->rho = rho;
    bool w = false;
    const Geom gl = split_geom(h, adj_tile, &w);
    const bool lam = w && !gl.nf;
    const AdjointLayout L = adjoint_layout(h->g, n, h->num_cus, adjoint_occupancy(w, lam, grad_amp != nullptr));
    float4* fpart = nullptr;
    if (fz != nullptr) {
        // the forward's partials after the adjoint's scratch
        const size_t off = (L.bytes + 255) / 256 * 256;
        const size_t fbytes = (size_t)fz->ch.split * h->g.n0 * h->g.n1 * sizeof(float4);
        if (int rc = ensure_scratch(h, off + fbytes)) return rc;
        fpart = reinterpret_cast<float4*>(static_cast<char*>(h->scratch) + off);
        if (int rc = run_forward_partials(h, fz->c, fz->ch.tile, fz->ch.split, fpart)) return rc;
    } else if (int rc = ensure_scratch(h, L.bytes)) {
        return rc;
    }
    IDISIM_HIP(launch_adjoint(w, lam, gl, L, emitters, amplitudes, n, static_cast<const float2*>(field), dLdI,
                              target, loss, h->scratch, grad_pos, grad_phase, grad_amp, st, fpart,
                              fz ? fz->ch.split : 0, fz ? static_cast<float2*>(const_cast<void*>(field)) : nullptr));
    const int splits = L.splits;
    h->plan.precision = IDISIM_PREC_FP32SPLIT;
    h->plan.tile = adj_tile;
    h->plan.tile_rows = tile_rows(adj_tile);
    h->plan.tile_cols = tile_cols(adj_tile);
    h->plan.split = splits;
    h->plan.phase_bound = bound;
    h->plan.fallback = 0;
    h->plan.graph = 0;
    return IDISIM_OK;
}

int idisim_adjoint(idisim_handle* h, const double* emitters, const double* amplitudes, int64_t n, double extent,
                   const void* field, const float* dLdI, float* grad_pos, float* grad_phase, float* grad_amp,
                   void* stream)
{
    return adjoint_call("idisim_adjoint", h, emitters, amplitudes, n, extent, field, dLdI, nullptr, nullptr, grad_pos,
                        grad_phase, grad_amp, stream);
}

int idisim_adjoint_l2(idisim_handle* h, const double* emitters, const double* amplitudes, int64_t n, double extent,
                      const void* field, const float* target, double* loss, float* grad_pos, float* grad_phase,
                      float* grad_amp, void* stream)
{
    return adjoint_call("idisim_adjoint_l2", h, emitters, amplitudes, n, extent, field, nullptr, target, loss,
                        grad_pos, grad_phase, grad_amp, stream);
}

int idisim_loss_grad_l2(idisim_handle* h, const double* emitters, const double* amplitudes, int64_t n, double extent,
                        const float* target, void* field, double* loss, float* grad_pos, float* grad_phase,
                        float* grad_amp, void* stream)
{
    const char* who = "idisim_loss_grad_l2";
    g_err.clear();
    if (!h) return fail(IDISIM_E_INVALID, std::string(who) + ": NULL handle");
    if (n < 1 || !emitters || !target || !grad_pos || !grad_phase)
        return fail(IDISIM_E_INVALID, std::string(who) + ": bad arguments (n >= 1, emitters, target and the gradients)");
    if (!std::isfinite(extent)) return fail(IDISIM_E_INVALID, std::string(who) + ": extent must be finite");
    if (h->precision != IDISIM_PREC_FP32SPLIT)
        return fail(IDISIM_E_UNSUPPORTED, std::string(who) + ": only FP32-split handles have an adjoint");
    DeviceGuard guard(h->device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (int rc = check_extent_flag(h, who)) return rc;
    double rho = extent;
    if (!(rho > 0.0)) {
        if (int rc = measure_extent(h, emitters, n, st, &rho)) return rc;
    }
    std::vector<int> tiles;
    bool split_path = true;
    if (int rc = choose_path(h, rho, &tiles, &split_path)) return rc;
    // the field: the caller's, or the handle's buffer (the autotuner and a forward without
    // split partials write one)
    void* fbuf = field;
    if (fbuf == nullptr) {
        if (!h->field_tmp) {
            const size_t fb = (size_t)h->g.n0 * h->g.n1 * sizeof(float2);
            if (hipMalloc(&h->field_tmp, fb) != hipSuccess) {
                h->field_tmp = nullptr;
                return fail(IDISIM_E_NOMEM, std::string(who) + ": field buffer");
            }
        }
        fbuf = h->field_tmp;
    }
    // an FP64 fallback forward writes the complex64 field through its finalize (out_f32)
    FusedForward fz{FwdCall{emitters, amplitudes, n, fbuf, IDISIM_OUT_FIELD, st, split_path, !split_path}, Choice{}};
    if (int rc = plan_choose(h, fz.c, tiles, fz.ch)) return rc;
    const bool fused = split_path && tile_pair(fz.ch.tile) && fz.ch.tile != kTileBlock && fz.ch.split > 1;
    if (!fused) {  // the forward with its own finalize, then idisim_adjoint_l2
        if (int rc = run_forward(h, fz.c, fz.ch.tile, fz.ch.split)) return rc;
        plan_record(h, fz.c, fz.ch);
        return adjoint_call(who, h, emitters, amplitudes, n, rho, fbuf, nullptr, target, loss, grad_pos, grad_phase,
                            grad_amp, stream);
    }
    plan_record(h, fz.c, fz.ch);
    return adjoint_call(who, h, emitters, amplitudes, n, rho, field, nullptr, target, loss, grad_pos, grad_phase,
                        grad_amp, stream, &fz);
}

}  // extern "C"
