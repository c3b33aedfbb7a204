// emu_rk.hip -- one stage combination of AMReX's Runge-Kutta integrator in
// one pass over the particles.
//
// The integrator the reference runs (main.cpp:129,202-213: AMReX
// TimeIntegrator with integration.rk.type) forms each stage state, and the
// step's result, as a copy of the old state followed by one saxpy
// (ApplyFlavoredNeutrinoRHS, FlavoredNeutrinoContainer.H:33-47) per earlier
// stage: p = S_old; p = (dt a_j) F_j + p for j = 0 .. n-1.  Done call by call
// that is 1 + n passes over every attribute; here it is one, with the same
// roundings in the same order (no contraction), so the result is bit for bit
// the copy-then-saxpys sequence.
#include "emu_internal.h"

#include <algorithm>

namespace emu {
namespace {

constexpr int kMaxStages = 4;

struct Stages {
    PView F[kMaxStages];
    double a[kMaxStages];
    int n;
};

__global__ __launch_bounds__(256) void k_rk_combine(PView dst, PView src, Stages st, int nattr)
{
#pragma clang fp contract(off)
    const int64_t ip = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (ip >= dst.np) return;
#pragma unroll
    for (int d = 0; d < 3; d++) dst.pos[d][ip] = src.pos[d][ip];
    for (int k = 0; k < nattr; k++) {
        double v = src.r[k][ip];
        for (int j = 0; j < st.n; j++) v = st.a[j] * st.F[j].r[k][ip] + v;
        dst.r[k][ip] = v;
    }
}

bool view_of(const emu_ctx* ctx, const emu_particles* p, PView& v)
{
    if (!p || p->np < 0) return false;
    v.np = p->np;
    const int na = 12 + 2 * ctx->p.nf * ctx->p.nf;
    for (int d = 0; d < 3; d++) {
        v.pos[d] = p->pos[d];
        if (p->np > 0 && !v.pos[d]) return false;
    }
    for (int k = 0; k < EMU_MAX_ATTRIBS; k++) {
        v.r[k] = k < na ? p->rdata[k] : nullptr;
        if (k < na && p->np > 0 && !v.r[k]) return false;
    }
    return true;
}

} // namespace
} // namespace emu

using namespace emu;

extern "C" int emu_rk_combine(emu_ctx* ctx, emu_particles* dst, const emu_particles* src, int n, const double* a,
                              const emu_particles* F)
{
    if (!ctx || n < 0 || n > kMaxStages || (n > 0 && (!a || !F))) return EMU_ERR_ARG;
    PView D, S;
    if (!view_of(ctx, dst, D) || !view_of(ctx, src, S) || D.np != S.np)
        return set_error(ctx, EMU_ERR_ARG, "bad particles");
    Stages st{};
    st.n = n;
    for (int j = 0; j < n; j++) {
        if (!view_of(ctx, &F[j], st.F[j]) || st.F[j].np != D.np) return set_error(ctx, EMU_ERR_ARG, "bad stage particles");
        st.a[j] = a[j];
    }
    const int na = 12 + 2 * ctx->p.nf * ctx->p.nf;
    if (D.np > 0) {
        const unsigned nb = (unsigned)std::max<int64_t>(1, (D.np + 255) / 256);
        hipLaunchKernelGGL(k_rk_combine, dim3(nb), dim3(256), 0, ctx->stream, D, S, st, na);
    }
    EMU_LAUNCH_CHECK(ctx, "rk_combine");
    return EMU_OK;
}
