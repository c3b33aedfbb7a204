// emu_ref.hip -- the per-function entry points: the reference's functions on
// the reference's data (SoA particles in PIdx order, the MultiFab `state` as
// one ghost-cell fab).  Built with -ffp-contract=off so that every product and
// sum here rounds as the reference writes it.  The deposit and the gather
// themselves run over particles listed by stencil box (emu_cells.hip); the
// deposit adds to the fab with FP64 atomics, as the reference's GPU build does
// (Gpu::Atomic::AddNoRet, generate_code.py:154), so its summation order is
// not deterministic.
#include "emu_internal.h"
#include "emu_cells.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

namespace emu {

int set_hip_error(emu_ctx* ctx, hipError_t e, const char* what)
{
    if (ctx) ctx->err = std::string(what) + ": " + hipGetErrorString(e);
    return EMU_ERR_HIP;
}

int set_error(emu_ctx* ctx, int code, const std::string& what)
{
    if (ctx) ctx->err = what;
    return code;
}

namespace {

constexpr int kBlock = 256;

inline unsigned nblocks(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kBlock - 1) / kBlock); }

__global__ void k_zero_comps(Fab fab, int c0, int c1)
{
    const int64_t n = fab.cstride * (c1 - c0);
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        fab.d[(int64_t)c0 * fab.cstride + t] = 0.0;
}

// SumBoundary: every valid cell adds its periodic ghost images.
__global__ void k_sum_boundary(Fab fab, Geom g, int ng0, int ng1, int ng2, int c0, int c1)
{
    const int64_t ncell = (int64_t)g.n[0] * g.n[1] * g.n[2];
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= ncell) return;
    const int i = (int)(t % g.n[0]), j = (int)((t / g.n[0]) % g.n[1]), k = (int)(t / ((int64_t)g.n[0] * g.n[1]));
    const int ng[3] = {ng0, ng1, ng2};
    for (int c = c0; c < c1; c++) {
        double s = fab.d[fab.at(i, j, k, c)];
        for (int sk = -1; sk <= 1; sk++)
            for (int sj = -1; sj <= 1; sj++)
                for (int si = -1; si <= 1; si++) {
                    if (si == 0 && sj == 0 && sk == 0) continue;
                    const int ii = i + si * g.n[0], jj = j + sj * g.n[1], kk = k + sk * g.n[2];
                    if (ii < -ng[0] || ii >= g.n[0] + ng[0] || jj < -ng[1] || jj >= g.n[1] + ng[1] ||
                        kk < -ng[2] || kk >= g.n[2] + ng[2])
                        continue;
                    s += fab.d[fab.at(ii, jj, kk, c)];
                }
        fab.d[fab.at(i, j, k, c)] = s;
    }
}

// FillBoundary: ghosts take their valid images (all components).
__global__ void k_fill_boundary(Fab fab, Geom g, int ncomp)
{
    const int64_t nfab = fab.cstride;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nfab) return;
    const int i = fab.lo[0] + (int)(t % fab.nf[0]);
    const int j = fab.lo[1] + (int)((t / fab.nf[0]) % fab.nf[1]);
    const int k = fab.lo[2] + (int)(t / ((int64_t)fab.nf[0] * fab.nf[1]));
    if (i >= 0 && i < g.n[0] && j >= 0 && j < g.n[1] && k >= 0 && k < g.n[2]) return;
    const int si = pmod(i, g.n[0]), sj = pmod(j, g.n[1]), sk = pmod(k, g.n[2]);
    for (int c = 0; c < ncomp; c++) fab.d[fab.at(i, j, k, c)] = fab.d[fab.at(si, sj, sk, c)];
}

__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* a, double v)
{
    // Non-negative doubles order like their bit patterns.
    atomicMax(a, (unsigned long long)__double_as_longlong(v > 0.0 ? v : 0.0));
}

// compute_dt's reduction (Evolve.cpp:40-56, fill generate_code.py:257-294).
template <int NF>
__global__ void k_compute_dt(Fab fab, Geom g, unsigned long long* red)
{
#pragma clang fp contract(off)
    using G = GIdx<NF>;
    constexpr int nh = NH<NF>;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    double Vad = 0.0, Vst = 0.0;
    if (t < fab.cstride) {
        const double* c0 = fab.d + t;
#define F(comp) c0[(int64_t)(comp) * fab.cstride]
        double hsi[nh], fl[nh];
        const double matter = F(G::rho) * F(G::Ye) / phys::Mp * g.cell_volume;
#pragma unroll
        for (int c = 0; c < nh; c++) hsi[c] = F(G::var(0, 0) + c) - F(G::var(0, 1) + c);
#pragma unroll
        for (int a = 0; a < NF; a++)
#pragma unroll
            for (int b = a + 1; b < NF; b++) {
                const int ci = herm_im<NF>(a, b);
                hsi[ci] = F(G::var(0, 0) + ci) + F(G::var(0, 1) + ci); // N - conj(Nbar)
            }
        hsi[0] += matter;
        double V2 = su_mag2<NF, double>(hsi);
#pragma unroll
        for (int dd = 1; dd <= 3; dd++) {
#pragma unroll
            for (int c = 0; c < nh; c++) fl[c] = F(G::var(dd, 0) + c) - F(G::var(dd, 1) + c);
            V2 += su_mag2<NF, double>(fl);
        }
        Vad = sqrt(V2) * M_SQRT2 * phys::GF / g.cell_volume;
#pragma unroll
        for (int a = 0; a < NF; a++) {
            const double x = F(G::var(0, 0) + herm_re<NF>(a, a)), y = F(G::var(0, 1) + herm_re<NF>(a, a));
            Vst = Vst > x ? Vst : x;
            Vst = Vst > y ? Vst : y;
        }
        Vst = Vst > matter ? Vst : matter;
        Vst *= 2.0 * NF * M_SQRT2 * phys::GF / g.cell_volume;
#undef F
    }
    // wave-level max, then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        Vad = fmax(Vad, __shfl_xor(Vad, off));
        Vst = fmax(Vst, __shfl_xor(Vst, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomic_max_nonneg(&red[0], Vad);
        atomic_max_nonneg(&red[1], Vst);
    }
}

__global__ void k_fill_u64(unsigned long long* p, int n, unsigned long long v)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = v;
}

// grid-stride over at most 1024 workgroups, so that one word takes a few
// thousand atomics rather than one per wave of particles
__global__ void k_min_pupt(const double* __restrict__ pupt, int64_t np, unsigned long long* red)
{
    double v = INFINITY;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < np; t += (int64_t)gridDim.x * blockDim.x)
        v = fmin(v, pupt[t]);
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0) atomicMin(red, (unsigned long long)__double_as_longlong(v));
}

// Renormalize (FlavoredNeutrinoContainer.cpp:107-111, fill generate_code.py:394-445).
template <int NF>
__device__ __forceinline__ int renorm_one(double* f, double L, double maxError)
{
#pragma clang fp contract(off)
    double sumP = 0;
#pragma unroll
    for (int i = 0; i < NF; i++) sumP += f[herm_re<NF>(i, i)];
    double error = sumP - 1.0;
    if (fabs(error) > 100. * maxError) return EMU_ERR_TRACE;
    if (fabs(error) > maxError) {
#pragma unroll
        for (int i = 0; i < NF; i++) f[herm_re<NF>(i, i)] -= error / NF;
    }
#pragma unroll
    for (int i = 0; i < NF; i++) {
        double& fii = f[herm_re<NF>(i, i)];
        if (fii < -100. * maxError) return EMU_ERR_NEGDIAG;
        if (fii < -maxError) fii = 0;
    }
    const double length = sqrt(su_mag2<NF, double>(f));
    error = length - L;
    if (fabs(error) > 100. * maxError) return EMU_ERR_LENGTH;
    if (fabs(error) > maxError) {
#pragma unroll
        for (int c = 0; c < NH<NF>; c++) f[c] /= length / L;
    }
    return EMU_OK;
}

template <int NF>
__global__ void k_renormalize(PView P, double maxError, unsigned long long* red)
{
    using I = PIdx<NF>;
    constexpr int nh = NH<NF>;
    const int64_t ip = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (ip >= P.np) return;
    int err = EMU_OK;
#pragma unroll
    for (int bar = 0; bar < 2 && err == EMU_OK; bar++) {
        double f[nh];
        const int f0 = bar ? I::fbar : I::f;
#pragma unroll
        for (int c = 0; c < nh; c++) f[c] = P.r[f0 + c][ip];
        err = renorm_one<NF>(f, P.r[bar ? I::Lbar : I::L][ip], maxError);
        if (err == EMU_OK) {
#pragma unroll
            for (int c = 0; c < nh; c++) P.r[f0 + c][ip] = f[c];
        }
    }
    if (err != EMU_OK) atomicMin(red, status_word(ip, err));
}

__global__ void k_sync(PView P, int type, int x0)
{
    const int64_t ip = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (ip >= P.np) return;
    for (int d = 0; d < 3; d++) {
        if (type == 0) P.pos[d][ip] = P.r[x0 + d][ip];
        else P.r[x0 + d][ip] = P.pos[d][ip];
    }
}

__device__ __forceinline__ double wrap1(double x, double rlo, double rhi, double L)
{
#pragma clang fp contract(off)
    // one period per step at most; the bound keeps non-finite input from looping
    if (x > rhi) {
        for (int it = 0; it < 8 && x > rhi; it++) x -= L;
        if (x < rlo || x > rhi) x = rlo;
    } else if (x < rlo) {
        for (int it = 0; it < 8 && x < rlo; it++) x += L;
        if (x > rhi || x < rlo) x = rhi;
    }
    return x;
}

__global__ void k_wrap(PView P, Geom g)
{
    const int64_t ip = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (ip >= P.np) return;
    for (int d = 0; d < 3; d++) P.pos[d][ip] = wrap1(P.pos[d][ip], g.rlo[d], g.rhi[d], g.L[d]);
}

__global__ void k_apply_rhs(PView P, double a, PView dF, int nattr)
{
#pragma clang fp contract(off)
    const int64_t ip = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (ip >= P.np) return;
    for (int k = 0; k < nattr; k++) P.r[k][ip] = a * dF.r[k][ip] + P.r[k][ip];
}

__global__ void k_copy(PView dst, PView src, int nattr)
{
    const int64_t ip = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (ip >= dst.np) return;
    for (int d = 0; d < 3; d++) dst.pos[d][ip] = src.pos[d][ip];
    for (int k = 0; k < nattr; k++) dst.r[k][ip] = src.r[k][ip];
}

} // namespace

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
static void mass_matrix(const emu_params& p, double* M2)
{
    // M2 = U diag(m^2) U^dagger, U = U12 P (NF=2) or U23 U13 U12 P (NF=3) with
    // U23 built from theta13 in the (0,2) plane (generate_code.py:176-227).
    double ur[3][3] = {{std::cos(p.theta12), std::sin(p.theta12), 0},
                       {-std::sin(p.theta12), std::cos(p.theta12), 0},
                       {0, 0, 1}};
    double ui[3][3] = {{0}};
    if (p.nf == 3) {
        const double c13 = std::cos(p.theta13), s13 = std::sin(p.theta13);
        const double cd = std::cos(p.deltaCP), sd = std::sin(p.deltaCP);
        const double br[3][3] = {{c13, 0, s13 * cd}, {0, 1, 0}, {-s13 * cd, 0, c13}};
        const double bi[3][3] = {{0, 0, -s13 * sd}, {0, 0, 0}, {-s13 * sd, 0, 0}};
        const double cr[3][3] = {{c13, 0, s13}, {0, 1, 0}, {-s13, 0, c13}};
        double tr[3][3], ti[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double sr = 0, si = 0;
                for (int k = 0; k < 3; k++) {
                    sr += br[i][k] * ur[k][j] - bi[i][k] * ui[k][j];
                    si += br[i][k] * ui[k][j] + bi[i][k] * ur[k][j];
                }
                tr[i][j] = sr;
                ti[i][j] = si;
            }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double sr = 0, si = 0;
                for (int k = 0; k < 3; k++) {
                    sr += cr[i][k] * tr[k][j];
                    si += cr[i][k] * ti[k][j];
                }
                ur[i][j] = sr;
                ui[i][j] = si;
            }
    }
    const double m2[3] = {p.mass[0] * p.mass[0], p.mass[1] * p.mass[1], p.mass[2] * p.mass[2]};
    int n = 0;
    for (int i = 0; i < p.nf; i++)
        for (int j = i; j < p.nf; j++) {
            double sr = 0, si = 0;
            for (int k = 0; k < p.nf; k++) {
                sr += m2[k] * (ur[i][k] * ur[j][k] + ui[i][k] * ui[j][k]);
                si += m2[k] * (ui[i][k] * ur[j][k] - ur[i][k] * ui[j][k]);
            }
            M2[n++] = sr;
            if (j > i) M2[n++] = si;
        }
}

static bool make_fab(const emu_ctx* ctx, const emu_grid* gr, Fab& fab)
{
    if (!gr || !gr->data) return false;
    const int nc = 3 + 8 * ctx->p.nf * ctx->p.nf;
    if (gr->ncomp != nc) return false;
    fab.d = gr->data;
    for (int d = 0; d < 3; d++) {
        fab.lo[d] = gr->lo[d];
        fab.nf[d] = gr->hi[d] - gr->lo[d] + 1;
        // the fab must hold the ghost cells the stencil reaches
        if (gr->lo[d] > -ctx->ngrow[d] || gr->hi[d] < ctx->p.ncell[d] - 1 + ctx->ngrow[d]) return false;
    }
    fab.cstride = (int64_t)fab.nf[0] * fab.nf[1] * fab.nf[2];
    return true;
}

static bool make_pview(const emu_ctx* ctx, const emu_particles* p, PView& v, bool need_pos = true)
{
    if (!p || p->np < 0) return false;
    v.np = p->np;
    const int na = 12 + 2 * ctx->p.nf * ctx->p.nf;
    for (int d = 0; d < 3; d++) {
        v.pos[d] = p->pos[d];
        if (need_pos && p->np > 0 && !v.pos[d]) return false;
    }
    for (int k = 0; k < EMU_MAX_ATTRIBS; k++) {
        v.r[k] = k < na ? p->rdata[k] : nullptr;
        if (k < na && p->np > 0 && !v.r[k]) return false;
    }
    return true;
}

} // namespace emu

using namespace emu;

extern "C" int emu_destroy(emu_ctx* ctx);

namespace emu {
namespace {
// deposit_to_mesh (Evolve.cpp:91-126) into `fab` with the given box-list
// scratch: zero the deposited components (valid and ghost cells), deposit,
// then SumBoundary.  Launches kernels only.
int deposit_into(emu_ctx* ctx, const PView& P, const Fab& fab, void* scratch, size_t bytes)
{
    const int nc = 3 + 8 * ctx->p.nf * ctx->p.nf;
    hipLaunchKernelGGL(k_zero_comps, dim3(1024), dim3(kBlock), 0, ctx->stream, fab, 3, nc);
    const int e = box_list_and_run(ctx, P, fab, true, scratch, bytes);
    if (e != EMU_OK) return e;
    const int64_t ncell = (int64_t)ctx->g.n[0] * ctx->g.n[1] * ctx->g.n[2];
    hipLaunchKernelGGL(k_sum_boundary, dim3(nblocks(ncell)), dim3(kBlock), 0, ctx->stream, fab, ctx->g,
                       ctx->ngrow[0], ctx->ngrow[1], ctx->ngrow[2], 3, nc);
    EMU_LAUNCH_CHECK(ctx, "deposit_to_mesh");
    return EMU_OK;
}
} // namespace
} // namespace emu

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
extern "C" int emu_create(const emu_params* params, emu_ctx** out)
{
    if (!params || !out) return EMU_ERR_ARG;
    *out = nullptr;
    const emu_params& p = *params;
    if (p.nf != 2 && p.nf != 3) return EMU_ERR_UNSUPPORTED;
    if (p.shape_order < 0 || p.shape_order > 2) return EMU_ERR_UNSUPPORTED;
    auto* ctx = new emu_ctx();
    ctx->p = p;
    Geom& g = ctx->g;
    for (int d = 0; d < 3; d++) {
        if (p.ncell[d] < 1 || !(p.phi[d] > p.plo[d])) { delete ctx; return EMU_ERR_ARG; }
        g.n[d] = p.ncell[d];
        g.order[d] = p.ncell[d] > 1 ? p.shape_order : 0;
        ctx->ngrow[d] = 1 + (1 + g.order[d]) / 2;
        if (p.ncell[d] < ctx->ngrow[d] && g.order[d] > 0) { delete ctx; return EMU_ERR_UNSUPPORTED; } // main.cpp:72
        g.plo[d] = p.plo[d];
        g.L[d] = p.phi[d] - p.plo[d];
        g.dx[d] = g.L[d] / (double)p.ncell[d];
        g.dxi[d] = 1.0 / g.dx[d];
        g.rlo[d] = p.plo[d];
        double r = p.phi[d];
        while ((int)std::floor((r - p.plo[d]) * g.dxi[d]) > p.ncell[d] - 1) r = std::nextafter(r, -INFINITY);
        g.rhi[d] = r;
    }
    g.cell_volume = g.dx[0] * g.dx[1] * g.dx[2];
    g.inv_cell_volume = g.dxi[0] * g.dxi[1] * g.dxi[2];
    mass_matrix(p, ctx->M2);
    if (hipMalloc(&ctx->d_red, 8 * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&ctx->d_M2, 9 * sizeof(double)) != hipSuccess) {
        emu_destroy(ctx);
        return EMU_ERR_NOMEM;
    }
    // the static scratch of emu_deposit_to_mesh_static, looked up once per
    // device; if that fails, the static entry point reports it on use
    (void)static_scratch_init();
    if (hipMemset(ctx->d_red, 0, 8 * sizeof(unsigned long long)) != hipSuccess ||
        hipMemcpy(ctx->d_M2, ctx->M2, 9 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        emu_destroy(ctx);
        return EMU_ERR_HIP;
    }
    *out = ctx;
    return EMU_OK;
}

extern "C" int emu_destroy(emu_ctx* ctx)
{
    if (!ctx) return EMU_OK;
    if (ctx->d_red) (void)hipFree(ctx->d_red);
    if (ctx->d_M2) (void)hipFree(ctx->d_M2);
    if (ctx->box_scratch) (void)hipFree(ctx->box_scratch);
    delete ctx;
    return EMU_OK;
}

extern "C" int emu_set_stream(emu_ctx* ctx, void* stream)
{
    if (!ctx) return EMU_ERR_ARG;
    ctx->stream = (hipStream_t)stream;
    return EMU_OK;
}

extern "C" int emu_sync(emu_ctx* ctx)
{
    if (!ctx) return EMU_ERR_ARG;
    unsigned long long oob = 0;
    EMU_HIP(ctx, hipMemcpyAsync(&oob, ctx->d_red + 7, sizeof oob, hipMemcpyDeviceToHost, ctx->stream));
    EMU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (oob) {
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, ctx->d_red + 7, 1, 0ull);
        return set_error(ctx, EMU_ERR_ARG, std::to_string(oob) + " particle stencils fell outside the fab and were skipped");
    }
    return EMU_OK;
}

extern "C" const char* emu_last_error(const emu_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int emu_grid_ngrow(const emu_ctx* ctx, int ngrow[3])
{
    if (!ctx || !ngrow) return EMU_ERR_ARG;
    for (int d = 0; d < 3; d++) ngrow[d] = ctx->ngrow[d];
    return EMU_OK;
}

// ---------------------------------------------------------------------------
// per-function entry points
// ---------------------------------------------------------------------------
extern "C" int emu_deposit_to_mesh(emu_ctx* ctx, const emu_particles* parts, emu_grid* state)
{
    if (!ctx) return EMU_ERR_ARG;
    Fab fab;
    PView P;
    if (!make_fab(ctx, state, fab) || !make_pview(ctx, parts, P)) return set_error(ctx, EMU_ERR_ARG, "bad grid or particles");
    const size_t need = box_scratch_size(P.np, fab.cstride, ctx->p.nf, false).total;
    void* scratch = nullptr;
    const int e = box_scratch(ctx, need, &scratch);
    if (e != EMU_OK) return e;
    return deposit_into(ctx, P, fab, scratch, need);
}

extern "C" int emu_deposit_to_mesh_static(emu_ctx* ctx, const emu_particles* parts, emu_grid* state)
{
    if (!ctx) return EMU_ERR_ARG;
    Fab fab;
    PView P;
    if (!make_fab(ctx, state, fab) || !make_pview(ctx, parts, P)) return set_error(ctx, EMU_ERR_ARG, "bad grid or particles");
    void* scratch = nullptr;
    size_t bytes = 0;
    const int e = static_scratch(ctx, P.np, fab.cstride, &scratch, &bytes);
    if (e != EMU_OK) return e;
    return deposit_into(ctx, P, fab, scratch, bytes);
}

extern "C" int emu_fill_boundary(emu_ctx* ctx, emu_grid* state)
{
    if (!ctx) return EMU_ERR_ARG;
    Fab fab;
    if (!make_fab(ctx, state, fab)) return set_error(ctx, EMU_ERR_ARG, "bad grid");
    hipLaunchKernelGGL(k_fill_boundary, dim3(nblocks(fab.cstride)), dim3(kBlock), 0, ctx->stream, fab, ctx->g,
                       state->ncomp);
    EMU_LAUNCH_CHECK(ctx, "fill_boundary");
    return EMU_OK;
}

extern "C" int emu_interpolate_rhs_from_mesh(emu_ctx* ctx, emu_particles* rhs, const emu_grid* state)
{
    if (!ctx) return EMU_ERR_ARG;
    Fab fab;
    PView P;
    if (!make_fab(ctx, state, fab) || !make_pview(ctx, rhs, P)) return set_error(ctx, EMU_ERR_ARG, "bad grid or particles");
    if (P.np == 0) return EMU_OK;
    const size_t need = box_scratch_size(P.np, fab.cstride, ctx->p.nf, true).total;
    void* scratch = nullptr;
    const int e = box_scratch(ctx, need, &scratch);
    if (e != EMU_OK) return e;
    return box_list_and_run(ctx, P, fab, false, scratch, need);
}

extern "C" int emu_compute_dt(emu_ctx* ctx, const emu_grid* state, double Vvac_max, double* dt_out)
{
    if (!ctx || !dt_out) return EMU_ERR_ARG;
    Fab fab;
    if (!make_fab(ctx, state, fab)) return set_error(ctx, EMU_ERR_ARG, "bad grid");
    const emu_params& p = ctx->p;
    const Geom& g = ctx->g;
    double dt_translation = 0.0;
    if (p.cfl_factor > 0.0) dt_translation = std::min(std::min(g.dx[0], g.dx[1]), g.dx[2]) / phys::c * p.cfl_factor;
    double dt_flavor = 0.0;
    if (p.flavor_cfl_factor > 0.0) {
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, ctx->d_red, 2, 0ull);
        if (p.nf == 2) hipLaunchKernelGGL(k_compute_dt<2>, dim3(nblocks(fab.cstride)), dim3(kBlock), 0, ctx->stream, fab, g, ctx->d_red);
        else hipLaunchKernelGGL(k_compute_dt<3>, dim3(nblocks(fab.cstride)), dim3(kBlock), 0, ctx->stream, fab, g, ctx->d_red);
        EMU_LAUNCH_CHECK(ctx, "compute_dt");
        unsigned long long h[2];
        EMU_HIP(ctx, hipMemcpyAsync(h, ctx->d_red, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
        EMU_HIP(ctx, hipStreamSynchronize(ctx->stream));
        double vad, vst;
        std::memcpy(&vad, &h[0], 8);
        std::memcpy(&vst, &h[1], 8);
        const double Vmax_adaptive = vad + Vvac_max;
        const double Vmax_stupid = vst + Vvac_max;
        const double dt_flavor_adaptive = phys::hbar / Vmax_adaptive * p.flavor_cfl_factor;
        const double dt_flavor_stupid = phys::hbar / Vmax_stupid * p.flavor_cfl_factor;
        dt_flavor = dt_flavor_stupid;
        if (p.max_adaptive_speedup > 1) dt_flavor = std::min(dt_flavor_stupid * p.max_adaptive_speedup, dt_flavor_adaptive);
    }
    if (dt_translation != 0.0 && dt_flavor != 0.0) *dt_out = std::min(dt_translation, dt_flavor);
    else if (dt_translation != 0.0) *dt_out = dt_translation;
    else if (dt_flavor != 0.0) *dt_out = dt_flavor;
    else return set_error(ctx, EMU_ERR_DT, "Timestep selection failed, try using both cfl_factor and flavor_cfl_factor");
    return EMU_OK;
}

extern "C" int emu_vvac_max(emu_ctx* ctx, const emu_particles* parts, double* out)
{
    if (!ctx || !out) return EMU_ERR_ARG;
    PView P;
    if (!make_pview(ctx, parts, P, false)) return set_error(ctx, EMU_ERR_ARG, "bad particles");
    const int ipupt = PIdx<2>::pupt; // same index for every NF
    const double inf = INFINITY;
    unsigned long long inf_bits;
    std::memcpy(&inf_bits, &inf, 8);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, ctx->d_red, 1, inf_bits);
    if (P.np > 0)
        hipLaunchKernelGGL(k_min_pupt, dim3(std::min(nblocks(P.np), 1024u)), dim3(kBlock), 0, ctx->stream, P.r[ipupt],
                           P.np, ctx->d_red);
    EMU_LAUNCH_CHECK(ctx, "vvac_max");
    unsigned long long h;
    EMU_HIP(ctx, hipMemcpyAsync(&h, ctx->d_red, 8, hipMemcpyDeviceToHost, ctx->stream));
    EMU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double pupt_min;
    std::memcpy(&pupt_min, &h, 8);
    // |diag(m1^2, m2^2, m3^2)|_SU (HermitianUtils.py:102-120)
    double d[9] = {0};
    if (ctx->p.nf == 2) {
        d[herm_re<2>(0, 0)] = ctx->p.mass[0] * ctx->p.mass[0];
        d[herm_re<2>(1, 1)] = ctx->p.mass[1] * ctx->p.mass[1];
        *out = std::sqrt(su_mag2<2, double>(d)) * phys::c4 / pupt_min;
    } else {
        for (int i = 0; i < 3; i++) d[herm_re<3>(i, i)] = ctx->p.mass[i] * ctx->p.mass[i];
        *out = std::sqrt(su_mag2<3, double>(d)) * phys::c4 / pupt_min;
    }
    return EMU_OK;
}

extern "C" int emu_renormalize(emu_ctx* ctx, emu_particles* parts, int64_t* bad_index)
{
    if (!ctx) return EMU_ERR_ARG;
    PView P;
    if (!make_pview(ctx, parts, P, false)) return set_error(ctx, EMU_ERR_ARG, "bad particles");
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, ctx->d_red, 1, ~0ull);
    if (P.np > 0) {
        if (ctx->p.nf == 2) hipLaunchKernelGGL(k_renormalize<2>, dim3(nblocks(P.np)), dim3(kBlock), 0, ctx->stream, P, ctx->p.maxError, ctx->d_red);
        else hipLaunchKernelGGL(k_renormalize<3>, dim3(nblocks(P.np)), dim3(kBlock), 0, ctx->stream, P, ctx->p.maxError, ctx->d_red);
    }
    EMU_LAUNCH_CHECK(ctx, "renormalize");
    unsigned long long h;
    EMU_HIP(ctx, hipMemcpyAsync(&h, ctx->d_red, 8, hipMemcpyDeviceToHost, ctx->stream));
    EMU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return status_decode(h, bad_index);
}

extern "C" int emu_sync_location(emu_ctx* ctx, emu_particles* parts, int type)
{
    if (!ctx || (type != 0 && type != 1)) return EMU_ERR_ARG;
    PView P;
    if (!make_pview(ctx, parts, P)) return set_error(ctx, EMU_ERR_ARG, "bad particles");
    if (P.np > 0) hipLaunchKernelGGL(k_sync, dim3(nblocks(P.np)), dim3(kBlock), 0, ctx->stream, P, type, (int)PIdx<2>::x);
    EMU_LAUNCH_CHECK(ctx, "sync_location");
    return EMU_OK;
}

extern "C" int emu_redistribute_local(emu_ctx* ctx, emu_particles* parts)
{
    if (!ctx) return EMU_ERR_ARG;
    PView P;
    if (!make_pview(ctx, parts, P)) return set_error(ctx, EMU_ERR_ARG, "bad particles");
    if (P.np > 0) hipLaunchKernelGGL(k_wrap, dim3(nblocks(P.np)), dim3(kBlock), 0, ctx->stream, P, ctx->g);
    EMU_LAUNCH_CHECK(ctx, "redistribute_local");
    return EMU_OK;
}

extern "C" int emu_apply_rhs(emu_ctx* ctx, emu_particles* parts, double a, const emu_particles* dFdt)
{
    if (!ctx) return EMU_ERR_ARG;
    PView P, D;
    if (!make_pview(ctx, parts, P, false) || !make_pview(ctx, dFdt, D, false) || D.np != P.np)
        return set_error(ctx, EMU_ERR_ARG, "bad particles");
    const int na = 12 + 2 * ctx->p.nf * ctx->p.nf;
    if (P.np > 0) hipLaunchKernelGGL(k_apply_rhs, dim3(nblocks(P.np)), dim3(kBlock), 0, ctx->stream, P, a, D, na);
    EMU_LAUNCH_CHECK(ctx, "apply_rhs");
    return EMU_OK;
}

extern "C" int emu_copy_particles(emu_ctx* ctx, emu_particles* dst, const emu_particles* src)
{
    if (!ctx) return EMU_ERR_ARG;
    PView Dv, Sv;
    if (!make_pview(ctx, dst, Dv) || !make_pview(ctx, src, Sv) || Dv.np != Sv.np)
        return set_error(ctx, EMU_ERR_ARG, "bad particles");
    const int na = 12 + 2 * ctx->p.nf * ctx->p.nf;
    if (Dv.np > 0) hipLaunchKernelGGL(k_copy, dim3(nblocks(Dv.np)), dim3(kBlock), 0, ctx->stream, Dv, Sv, na);
    EMU_LAUNCH_CHECK(ctx, "copy_particles");
    return EMU_OK;
}
