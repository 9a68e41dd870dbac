// Host build of the FP32-split error planner (idi_amd/csrc/planner.hpp) and of the C ABI's
// argument checks (idi_amd/csrc/checks.hpp) for the sanitizer test (tests/test_host_asan.py),
// compiled with a host compiler under AddressSanitizer and UndefinedBehaviorSanitizer.
//
// `planner_check checks` runs the argument checks on good and bad arguments and prints
// {"name": code, ...}.  `planner_check sincos` reads one double per line (stdin) and prints
// "%a %a", the sine and cosine of the FP64 kernels' sincos_fp64 (idi_amd/csrc/sincos64.hpp).
// Otherwise, input, one case per line (stdin):
//   n0 n1 pixelsize detz k off0 off1 nf scale secondorder rot m0 .. m8 tol_rev summode nrho rho_1 .. rho_nrho
// Output, one JSON object per case (stdout): the planner tables, and for every extent the
// bound of every tile variant (null where it is infinite) and the eligible variants.
#include "../../idi_amd/csrc/checks.hpp"
#include "../../idi_amd/csrc/planner.hpp"
#include "../../idi_amd/csrc/sincos64.hpp"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace idisim;

static void put(double v)
{
    if (std::isfinite(v))
        std::printf("%.17g", v);
    else
        std::printf("null");
}

static int checks()
{
    std::string m;
    idisim_geometry g{};
    g.n0 = 64;
    g.n1 = 48;
    g.pixelsize = 75.0;
    g.detz = 1e5;
    g.k = 4e4;
    double buf[64] = {};
    float fbuf[16] = {};
    const int64_t offs_ok[4] = {0, 3, 3, 7}, offs_bad0[3] = {1, 2, 3}, offs_dec[3] = {0, 5, 4};
    const int h = 1;  // a stand-in handle: the checks only test it against NULL
    struct Row {
        const char* name;
        int code;
    };
    std::vector<Row> r;
    auto add = [&](const char* name, int code) { r.push_back({name, code}); };
    add("create_ok", check_create(&g, IDISIM_PREC_FP32SPLIT, &m));
    add("create_null", check_create(nullptr, IDISIM_PREC_FP64, &m));
    add("create_precision", check_create(&g, 7, &m));
    idisim_geometry b = g;
    b.n0 = 0;
    add("create_empty", check_create(&b, IDISIM_PREC_FP64, &m));
    b = g;
    b.n0 = b.n1 = 1 << 20;
    add("create_too_many_pixels", check_create(&b, IDISIM_PREC_FP64, &m));
    b = g;
    b.k = -1.0;
    add("create_negative_k", check_create(&b, IDISIM_PREC_FP64, &m));
    b = g;
    b.detz = NAN;
    add("create_nan_detz", check_create(&b, IDISIM_PREC_FP64, &m));
    b = g;
    b.flags = IDISIM_ROTATE;
    b.rot[4] = INFINITY;
    add("create_inf_rot", check_create(&b, IDISIM_PREC_FP64, &m));
    add("forward_ok", check_forward("f", &h, buf, 16, buf, IDISIM_OUT_FIELD, 0.0, &m));
    add("forward_empty_ok", check_forward("f", &h, nullptr, 0, buf, IDISIM_OUT_INTENSITY, 0.0, &m));
    add("forward_null_handle", check_forward("f", nullptr, buf, 16, buf, IDISIM_OUT_FIELD, 0.0, &m));
    add("forward_null_out", check_forward("f", &h, buf, 16, nullptr, IDISIM_OUT_FIELD, 0.0, &m));
    add("forward_null_emitters", check_forward("f", &h, nullptr, 16, buf, IDISIM_OUT_FIELD, 0.0, &m));
    add("forward_negative_n", check_forward("f", &h, buf, -1, buf, IDISIM_OUT_FIELD, 0.0, &m));
    add("forward_kind", check_forward("f", &h, buf, 16, buf, 5, 0.0, &m));
    add("forward_nan_extent", check_forward("f", &h, buf, 16, buf, IDISIM_OUT_FIELD, NAN, &m));
    add("batched_ok", check_batched(&h, buf, offs_ok, 3, buf, IDISIM_OUT_FIELD, 0.1, &m));
    add("batched_offset0", check_batched(&h, buf, offs_bad0, 2, buf, IDISIM_OUT_FIELD, 0.1, &m));
    add("batched_decreasing", check_batched(&h, buf, offs_dec, 2, buf, IDISIM_OUT_FIELD, 0.1, &m));
    add("batched_no_shots", check_batched(&h, buf, offs_ok, 0, buf, IDISIM_OUT_FIELD, 0.1, &m));
    add("batched_null_emitters", check_batched(&h, nullptr, offs_ok, 3, buf, IDISIM_OUT_FIELD, 0.1, &m));
    add("adjoint_ok", check_adjoint("a", &h, buf, 4, buf, fbuf, fbuf, fbuf, 0.0, &m));
    add("adjoint_null_grad", check_adjoint("a", &h, buf, 4, buf, fbuf, nullptr, fbuf, 0.0, &m));
    add("adjoint_null_field", check_adjoint("a", &h, buf, 4, nullptr, fbuf, fbuf, fbuf, 0.0, &m));
    add("adjoint_null_target", check_adjoint("a", &h, buf, 4, buf, nullptr, fbuf, fbuf, 0.0, &m));
    add("time_ok", check_time_forward(&h, buf, buf, 4, 0.3, 1.0, buf, &m));
    add("time_tau", check_time_forward(&h, buf, buf, 4, 0.3, 0.0, buf, &m));
    add("option_tile_ok", check_option(IDISIM_OPT_TILE, kTileRefS, kNumTiles, &m));
    add("option_tile_reserved", check_option(IDISIM_OPT_TILE, kTileNone, kNumTiles, &m));
    add("option_tile_range", check_option(IDISIM_OPT_TILE, kNumTiles, kNumTiles, &m));
    add("option_unknown", check_option(999, 1, kNumTiles, &m));
    add("option_tol", check_option(IDISIM_OPT_TOL_PPB, 0, kNumTiles, &m));
    add("generate_radius", check_generate("g", 4, buf, -1.0, &m));
    add("generate_ok", check_generate("g", 4, buf, 1.0, &m));
    add("chargesharing_overlap", check_chargesharing(buf, 4, 4, 0.5, 0.0, buf + 3, &m));
    add("chargesharing_ok", check_chargesharing(buf, 4, 4, 0.5, 0.0, buf + 16, &m));
    add("chargesharing_sigma", check_chargesharing(buf, 4, 4, 40.0, 0.0, buf + 16, &m));
    add("qcorr_ok", check_qcorr(fbuf, fbuf, 4, 4, 1.5, 1.5, 100.0, 2, 1, buf, &m));
    add("qcorr_qmax", check_qcorr(fbuf, fbuf, 4, 4, 1.5, 1.5, 100.0, 2000, 1, buf, &m));
    std::printf("{");
    for (size_t i = 0; i < r.size(); ++i) std::printf("%s\"%s\": %d", i ? ", " : "", r[i].name, r[i].code);
    std::printf("}\n");
    return 0;
}

int main(int argc, char** argv)
{
    if (argc > 1 && std::strcmp(argv[1], "checks") == 0) return checks();
    if (argc > 1 && std::strcmp(argv[1], "sincos") == 0) {
        double x;
        while (std::scanf("%lf", &x) == 1) {
            double sn, cs;
            sincos_fp64(x, &sn, &cs);
            std::printf("%a %a\n", sn, cs);
        }
        return 0;
    }
    for (;;) {
        long n0, n1;
        double pix, detz, k, off0, off1, tol;
        int nf, scale, secondorder, rot, summode, nrho;
        double M[9];
        if (std::scanf("%ld %ld %lf %lf %lf %lf %lf %d %d %d %d", &n0, &n1, &pix, &detz, &k, &off0, &off1, &nf, &scale,
                       &secondorder, &rot) != 11)
            break;
        for (double& m : M)
            if (std::scanf("%lf", &m) != 1) return 2;
        if (std::scanf("%lf %d %d", &tol, &summode, &nrho) != 3 || nrho < 0 || nrho > 64) return 2;
        std::vector<double> rhos((size_t)nrho);
        for (double& r : rhos)
            if (std::scanf("%lf", &r) != 1) return 2;
        idisim_geometry geom{};
        geom.n0 = n0;
        geom.n1 = n1;
        geom.pixelsize = pix;
        geom.detz = detz;
        geom.k = k;
        geom.detoffset[0] = off0;
        geom.detoffset[1] = off1;
        for (int i = 0; i < 9; ++i) geom.rot[i] = M[i];
        geom.flags = (nf ? IDISIM_NF : 0u) | (scale ? IDISIM_SCALE : 0u) | (secondorder ? IDISIM_SECONDORDER : 0u) |
                     (rot ? IDISIM_ROTATE : 0u);
        std::string msg;
        if (check_create(&geom, IDISIM_PREC_FP32SPLIT, &msg) != IDISIM_OK) {
            std::printf("{\"error\": \"%s\"}\n", msg.c_str());
            continue;
        }
        Planner p;
        p.g = make_geom(geom);
        p.tol_rev = tol;
        p.summode = summode;
        planner_tables(&p);
        std::printf("{\"Rmin\": ");
        put(p.Rmin);
        std::printf(", \"Rmax\": ");
        put(p.Rmax);
        std::printf(", \"pair_coef\": ");
        put(p.pair_coef);
        std::printf(", \"block_coef\": ");
        put(p.block_coef);
        std::printf(", \"ref_coef\": ");
        put(p.ref_coef);
        std::printf(", \"refs_coef\": ");
        put(p.refs_coef);
        std::printf(", \"cent_coef\": ");
        put(p.cent_coef);
        std::printf(", \"lin_coef\": ");
        put(p.lin_coef);
        std::printf(", \"du_max\": [");
        for (int v = 0; v < kNumTiles; ++v) {
            if (v) std::printf(", ");
            put(p.du_max[v]);
        }
        std::printf("], \"extents\": [");
        for (size_t i = 0; i < rhos.size(); ++i) {
            if (i) std::printf(", ");
            std::printf("{\"rho\": %.17g, \"bound\": [", rhos[i]);
            for (int v = 0; v < kNumTiles; ++v) {
                if (v) std::printf(", ");
                put(plan_bound(&p, v, rhos[i]));
            }
            std::printf("], \"ref_angle\": [");
            for (int v = 0; v < kNumTiles; ++v) {
                if (v) std::printf(", ");
                put(ref_angle_bound(&p, v, rhos[i]));
            }
            std::printf("], \"ref_centre\": [");
            for (int v = 0; v < kNumTiles; ++v) {
                if (v) std::printf(", ");
                put(ref_trunc_centre(&p, v, rhos[i]));
            }
            double best = 0.0;
            const std::vector<int> ok = eligible_tiles(&p, rhos[i], &best);
            std::printf("], \"eligible\": [");
            for (size_t j = 0; j < ok.size(); ++j) std::printf(j ? ", %d" : "%d", ok[j]);
            std::printf("], \"best\": ");
            put(best);
            std::printf("}");
        }
        std::printf("]}\n");
        std::fflush(stdout);
    }
    return 0;
}
