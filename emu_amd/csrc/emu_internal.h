// emu_internal.h -- context layout and helpers shared by the C ABI sources.
#pragma once

#include "../../include/emu_amd.h"
#include "emu_common.h"

#include <hip/hip_runtime.h>
#include <string>

namespace emu {

// Geometry passed to kernels by value (Geometry of main.cpp:58-62: dx = L/n,
// dxi = 1/dx; rhi = AMReX's roundoff upper bound used by the periodic shift).
struct Geom {
    int n[3];
    int order[3]; // shape order per dim, 0 in 1-cell dims (Evolve.cpp:102-104)
    double plo[3], L[3], dx[3], dxi[3], rlo[3], rhi[3];
    double cell_volume, inv_cell_volume;
};

// The reference's MultiFab as one ghost-cell fab (Array4 layout).
struct Fab {
    double* d;
    int lo[3], nf[3];
    int64_t cstride;
    EMU_HD int64_t at(int i, int j, int k, int c) const
    {
        return ((int64_t)c * nf[2] + (k - lo[2])) * (int64_t)(nf[1] * nf[0]) +
               (int64_t)(j - lo[1]) * nf[0] + (i - lo[0]);
    }
};

// SoA particle view (PIdx order).
struct PView {
    double* pos[3];
    double* r[EMU_MAX_ATTRIBS];
    int64_t np;
};

// Deferred status words.  A kernel that finds an error folds
// (index << kStatusBits) | code into one 64-bit word with atomicMin, so the
// word ends holding the lowest offending index (its code is that particle's).
// ~0 means no error; an all-ones index means an error that names no particle
// (compute_dt), which any error with an index outranks.
constexpr int kStatusBits = 4;
static_assert(EMU_ERR_MOVE < (1 << kStatusBits), "status codes must fit the status field");
EMU_HD unsigned long long status_word(int64_t index, int code)
{
    return ((unsigned long long)index << kStatusBits) | (unsigned long long)code;
}
EMU_HD unsigned long long status_word_noindex(int code)
{
    return (~0ull << kStatusBits) | (unsigned long long)code;
}
inline int status_decode(unsigned long long w, int64_t* bad_index)
{
    if (w == ~0ull) {
        if (bad_index) *bad_index = -1;
        return EMU_OK;
    }
    const unsigned long long ix = w >> kStatusBits;
    if (bad_index) *bad_index = ix == (~0ull >> kStatusBits) ? -1 : (int64_t)ix;
    return (int)(w & ((1ull << kStatusBits) - 1));
}

} // namespace emu

struct emu_ctx {
    emu_params p;
    emu::Geom g;
    double M2[9];      // flavour-basis mass matrix, real components, grams^2
    double* d_M2 = nullptr; // device copy of M2
    int ngrow[3];
    hipStream_t stream = nullptr;
    unsigned long long* d_red = nullptr; // reduction scratch (8 slots)
    void* box_scratch = nullptr;         // per-function deposit/gather: box list and combined fields
    size_t box_scratch_bytes = 0;
    bool adj_status_armed = false;       // d_red[5] holds the adjoint steps' deferred status
    std::string err;
};

namespace emu {
int set_hip_error(emu_ctx* ctx, hipError_t e, const char* what);
int set_error(emu_ctx* ctx, int code, const std::string& what);
}

#define EMU_HIP(ctx, call)                                                                     \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) return emu::set_hip_error((ctx), e_, #call);                     \
    } while (0)

#define EMU_LAUNCH_CHECK(ctx, what)                                                            \
    do {                                                                                       \
        hipError_t e_ = hipGetLastError();                                                     \
        if (e_ != hipSuccess) return emu::set_hip_error((ctx), e_, what);                      \
    } while (0)
