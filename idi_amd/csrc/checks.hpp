// Argument checks of the C ABI entry points (idisim.hip) and the geometry set-up of a handle:
// host functions with no device call, so that a host build under sanitizers exercises them
// without a GPU.  Each check returns IDISIM_OK, or an error code with its message in *msg.
#pragma once
#include "../../include/idisim.h"
#include "geometry.hpp"

#include <cmath>
#include <cstdint>
#include <string>

namespace idisim {

inline int arg_error(std::string* msg, const char* who, const std::string& text)
{
    *msg = std::string(who) + ": " + text;
    return IDISIM_E_INVALID;
}

// The handle's FP64 geometry (the set-up of cpu.py:30-40 and cuda.py:53-74): pixel (x, y) sits at
// ((x - n0/2) pix + off0, (y - n1/2) pix + off1, detz), rotated by rot[] with IDISIM_ROTATE.
inline Geom make_geom(const idisim_geometry& geom)
{
    Geom g{};
    g.pix = geom.pixelsize;
    g.detz = geom.detz;
    g.k = geom.k;
    g.krev = geom.k * kInv2Pi;
    g.off0 = geom.detoffset[0];
    g.off1 = geom.detoffset[1];
    g.half0 = (double)geom.n0 / 2.0;
    g.half1 = (double)geom.n1 / 2.0;
    for (int i = 0; i < 9; ++i) g.M[i] = geom.rot[i];
    g.n0 = (int)geom.n0;
    g.n1 = (int)geom.n1;
    g.rot = (geom.flags & IDISIM_ROTATE) ? 1 : 0;
    g.nf = (geom.flags & IDISIM_NF) ? 1 : 0;
    g.scale = (geom.flags & IDISIM_SCALE) ? 1 : 0;
    g.secondorder = (geom.flags & IDISIM_SECONDORDER) ? 1 : 0;
    g.refrho = 0.f;
    g.refc = 0.f;
    return g;
}

inline int check_create(const idisim_geometry* geom, int precision, std::string* msg)
{
    const char* who = "idisim_create";
    if (!geom) return arg_error(msg, who, "NULL argument");
    if (geom->n0 <= 0 || geom->n1 <= 0 || geom->n0 > (1 << 20) || geom->n1 > (1 << 20) ||
        geom->n0 * geom->n1 > (int64_t)1 << 31)
        return arg_error(msg, who, "detector size out of range");
    if (!(geom->k > 0.0) || !std::isfinite(geom->k) || !std::isfinite(geom->pixelsize) || !std::isfinite(geom->detz))
        return arg_error(msg, who, "k, pixelsize and detz must be finite, k > 0");
    if ((geom->flags & IDISIM_ROTATE) != 0)
        for (int i = 0; i < 9; ++i)
            if (!std::isfinite(geom->rot[i])) return arg_error(msg, who, "rot must be finite");
    if (!std::isfinite(geom->detoffset[0]) || !std::isfinite(geom->detoffset[1]))
        return arg_error(msg, who, "detoffset must be finite");
    if (precision != IDISIM_PREC_FP64 && precision != IDISIM_PREC_FP32SPLIT)
        return arg_error(msg, who, "unknown precision");
    return IDISIM_OK;
}

// idisim_forward and idisim_forward_noalloc
inline int check_forward(const char* who, const void* h, const double* emitters, int64_t n, const void* out,
                         int out_kind, double extent, std::string* msg)
{
    if (!h) return arg_error(msg, who, "NULL handle");
    if (!out) return arg_error(msg, who, "NULL output");
    if (n < 0 || (n > 0 && !emitters)) return arg_error(msg, who, "bad emitter arguments");
    if (out_kind != IDISIM_OUT_FIELD && out_kind != IDISIM_OUT_INTENSITY) return arg_error(msg, who, "unknown output kind");
    if (!std::isfinite(extent)) return arg_error(msg, who, "extent must be finite");
    return IDISIM_OK;
}

// offsets: host array of nshots + 1 non-decreasing counts starting at 0
inline int check_batched(const void* h, const double* emitters, const int64_t* offsets, int64_t nshots, const void* out,
                         int out_kind, double extent, std::string* msg)
{
    const char* who = "idisim_forward_batched";
    if (!h) return arg_error(msg, who, "NULL handle");
    if (!out || !offsets || nshots <= 0) return arg_error(msg, who, "bad arguments");
    if (out_kind != IDISIM_OUT_FIELD && out_kind != IDISIM_OUT_INTENSITY) return arg_error(msg, who, "unknown output kind");
    if (offsets[0] != 0) return arg_error(msg, who, "offsets[0] must be 0");
    for (int64_t s = 0; s < nshots; ++s)
        if (offsets[s + 1] < offsets[s]) return arg_error(msg, who, "offsets must be non-decreasing");
    if (offsets[nshots] > 0 && !emitters) return arg_error(msg, who, "NULL emitters");
    if (!std::isfinite(extent)) return arg_error(msg, who, "extent must be finite");
    return IDISIM_OK;
}

// idisim_adjoint (G = dL/dI) and idisim_adjoint_l2 (G = the target intensity)
inline int check_adjoint(const char* who, const void* h, const double* emitters, int64_t n, const void* field,
                         const float* G, const float* grad_pos, const float* grad_phase, double extent,
                         std::string* msg)
{
    if (!h) return arg_error(msg, who, "NULL handle");
    if (n < 0 || (n > 0 && (!emitters || !grad_pos || !grad_phase)) || !field || !G)
        return arg_error(msg, who, "bad arguments");
    if (!std::isfinite(extent)) return arg_error(msg, who, "extent must be finite");
    return IDISIM_OK;
}

inline int check_time_forward(const void* h, const double* emitters, const double* times, int64_t n, double c,
                              double tau, const double* out, std::string* msg)
{
    const char* who = "idisim_time_forward";
    if (!h) return arg_error(msg, who, "NULL handle");
    if (n < 0 || n >= ((int64_t)1 << 31) || (n > 0 && (!emitters || !times)) || !out)
        return arg_error(msg, who, "bad arguments");
    if (!(c > 0.0) || !std::isfinite(c) || !(tau > 0.0) || !std::isfinite(tau))
        return arg_error(msg, who, "c and tau must be finite and > 0");
    return IDISIM_OK;
}

// IDISIM_OPT_* values
inline int check_option(int option, int64_t value, int num_tiles, std::string* msg)
{
    const char* who = "idisim_set_option";
    switch (option) {
        case IDISIM_OPT_AUTOTUNE:
        case IDISIM_OPT_GRAPH:
        case IDISIM_OPT_FALLBACK: return IDISIM_OK;
        case IDISIM_OPT_TILE:
            if (value < -1 || value >= num_tiles) return arg_error(msg, who, "tile variant out of range");
            if (value == kTileNone) {
                *msg = std::string(who) + ": tile variant 4 has no kernel";
                return IDISIM_E_UNSUPPORTED;
            }
            return IDISIM_OK;
        case IDISIM_OPT_SPLIT:
            return (value < 0 || value > 4096) ? arg_error(msg, who, "split out of range") : IDISIM_OK;
        case IDISIM_OPT_SUMMODE:
            return (value != 0 && value != 1) ? arg_error(msg, who, "summode must be 0 or 1") : IDISIM_OK;
        case IDISIM_OPT_TOL_PPB:
            return (value <= 0) ? arg_error(msg, who, "tolerance must be positive") : IDISIM_OK;
        case IDISIM_OPT_TIME_PATH:
            return (value < 0 || value > 2) ? arg_error(msg, who, "IDISIM_OPT_TIME_PATH must be 0, 1 or 2") : IDISIM_OK;
        default: return arg_error(msg, who, "unknown option " + std::to_string(option));
    }
}

inline int check_generate(const char* who, int64_t n, const double* out, double radius, std::string* msg)
{
    if (n < 0 || (n > 0 && !out)) return arg_error(msg, who, "bad arguments");
    if (!(radius >= 0.0) || !std::isfinite(radius)) return arg_error(msg, who, "radius must be finite and >= 0");
    return IDISIM_OK;
}

inline int check_chargesharing(const double* counts, int64_t n0, int64_t n1, double sigma, double gain_noise,
                               const double* out, std::string* msg)
{
    const char* who = "idisim_chargesharing";
    if (n0 <= 0 || n1 <= 0 || n0 > (1 << 20) || n1 > (1 << 20) || n0 * n1 > ((int64_t)1 << 31))
        return arg_error(msg, who, "image size out of range");
    if (!counts || !out) return arg_error(msg, who, "NULL buffer");
    const uintptr_t bytes = (uintptr_t)(n0 * n1) * sizeof(double);
    const uintptr_t a = reinterpret_cast<uintptr_t>(counts), b = reinterpret_cast<uintptr_t>(out);
    if (a < b + bytes && b < a + bytes) return arg_error(msg, who, "out overlaps counts");
    if (sigma > 32.0) return arg_error(msg, who, "sigma is limited to 32 pixels");
    if (!std::isfinite(gain_noise) || gain_noise < 0.0) return arg_error(msg, who, "gain_noise must be finite and >= 0");
    return IDISIM_OK;
}

inline int check_qcorr(const float* zd, const float* img, int64_t nr, int64_t nc, double xcenter, double ycenter,
                       double z, int64_t qmax, int64_t maxdqz, const double* out, std::string* msg)
{
    const char* who = "idisim_qcorr";
    if (!zd || !img || !out) return arg_error(msg, who, "NULL buffer");
    if (nr < 1 || nc < 1 || nr > 65535 || nc > (1 << 20)) return arg_error(msg, who, "image size out of range");
    if (qmax < 1 || qmax > 1024 || maxdqz < 1 || maxdqz > (1 << 20))
        return arg_error(msg, who, "qmax must be in [1, 1024] and maxdqz >= 1");
    if (!std::isfinite(xcenter) || !std::isfinite(ycenter) || !std::isfinite(z))
        return arg_error(msg, who, "centre and z must be finite");
    return IDISIM_OK;
}

}  // namespace idisim
