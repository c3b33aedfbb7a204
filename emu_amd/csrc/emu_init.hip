// emu_init.hip -- initial conditions on the device: InitParticles
// (Source/FlavoredNeutrinoContainerInit.cpp:136-661), simulation types 0-5.
//
// One box, no tiling: particles are ordered by cell with z fastest (:222),
// then location in the cell, then direction (:294).  Perturbations come from
// Philox4x32-10 (Salmon et al., SC'11) keyed by (seed, particle index, draw)
// instead of the reference's wall-clock-seeded amrex::Random (main.cpp:238-240),
// so every run and every implementation sees the same arrays.
#include "emu_internal.h"
#include "../../include/emu_init.h"

#include <cmath>
#include <vector>

namespace emu {
namespace {

struct InitDev {
    int nf, st, nx, ny, nz, nppc[3], ndirs, nlocs;
    double plo[3], dx[3];
    double scale_fac, pupt, ndens, ndensbar, amp, amp3, k3, st4_fh[3], st4_fhb[3], st4_ff, st4_ffb;
    double Z[3], ff[3], fdir[3][3], n5[3]; // Minerbo: e, a, x
    uint64_t seed;
    const double* dirs; // [ndirs][3]
};

__device__ __forceinline__ void philox(uint32_t c[4], uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// Uniform [0,1): 53 bits from two 32-bit outputs of counter (pid, draw>>1).
__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t pid, int draw)
{
    uint32_t c[4] = {(uint32_t)pid, (uint32_t)(pid >> 32), (uint32_t)(draw >> 1), 0u};
    philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint32_t hi = (draw & 1) ? c[2] : c[0], lo = (draw & 1) ? c[3] : c[1];
    return ((double)(hi >> 5) * 67108864.0 + (double)(lo >> 6)) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ double symmetric_uniform(uint64_t seed, uint64_t pid, int draw)
{
    return 2. * (uniform01(seed, pid, draw) - 0.5); // FlavoredNeutrinoContainerInit.cpp:111-113
}

template <int NF>
__global__ void k_init(InitDev D, PView P, int64_t* ids)
{
    using I = PIdx<NF>;
    const int64_t np = P.np;
    for (int64_t pid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; pid < np; pid += (int64_t)gridDim.x * blockDim.x) {
        const int64_t per_cell = (int64_t)D.nlocs * D.ndirs;
        const int64_t cellid = pid / per_cell;
        const int iloc = (int)((pid / D.ndirs) % D.nlocs), idir = (int)(pid % D.ndirs);
        const int ix = (int)(cellid / ((int64_t)D.ny * D.nz)), iy = (int)((cellid / D.nz) % D.ny),
                  iz = (int)(cellid % D.nz);
        // get_position_unit_cell (:84-97)
        const int nyz = D.nppc[1] * D.nppc[2];
        const int ixp = iloc / nyz, iyp = (iloc % nyz) % D.nppc[1], izp = (iloc % nyz) / D.nppc[1];
        const double r0 = (0.5 + ixp) / D.nppc[0], r1 = (0.5 + iyp) / D.nppc[1], r2 = (0.5 + izp) / D.nppc[2];
        const double x = D.plo[0] + (ix + r0) * D.dx[0], y = D.plo[1] + (iy + r1) * D.dx[1],
                     z = D.plo[2] + (iz + r2) * D.dx[2];
        const double u[3] = {D.dirs[3 * idir], D.dirs[3 * idir + 1], D.dirs[3 * idir + 2]};
        double r[I::nattribs];
        for (int k = 0; k < I::nattribs; k++) r[k] = 0.0;
        r[I::x] = x;
        r[I::y] = y;
        r[I::z] = z;
        double* f = r + I::f;
        double* fb = r + I::fbar;
        r[I::pupt] = D.pupt;
        if (D.st <= 4) {
            f[herm_re<NF>(0, 0)] = 1.0;
            fb[herm_re<NF>(0, 0)] = 1.0;
        }
        if (D.st == 0) {
            r[I::N] = 1.0;
            r[I::Nbar] = 1.0;
        } else if (D.st == 1) {
            r[I::N] = D.ndens * D.scale_fac;
            r[I::Nbar] = D.ndens * D.scale_fac;
        } else if (D.st == 2) {
            r[I::N] = D.ndens * D.scale_fac * (1. + u[2]);
            r[I::Nbar] = D.ndens * D.scale_fac * (1. - u[2]);
        } else if (D.st == 3) {
            const double v = D.amp * sin(D.k3 * z);
            f[herm_re<NF>(0, 1)] = v;
            fb[herm_re<NF>(0, 1)] = v;
            r[I::N] = D.ndens * D.scale_fac * (1. + u[2]);
            r[I::Nbar] = D.ndens * D.scale_fac * (1. - u[2]);
        } else if (D.st == 4) {
            const double a = D.amp;
            f[herm_re<NF>(0, 1)] = a * symmetric_uniform(D.seed, pid, 0);
            f[herm_im<NF>(0, 1)] = a * symmetric_uniform(D.seed, pid, 1);
            fb[herm_re<NF>(0, 1)] = a * symmetric_uniform(D.seed, pid, 2);
            fb[herm_im<NF>(0, 1)] = a * symmetric_uniform(D.seed, pid, 3);
            if (NF == 3) {
                f[herm_re<NF>(0, 2)] = a * symmetric_uniform(D.seed, pid, 4);
                f[herm_im<NF>(0, 2)] = a * symmetric_uniform(D.seed, pid, 5);
                fb[herm_re<NF>(0, 2)] = a * symmetric_uniform(D.seed, pid, 6);
                fb[herm_im<NF>(0, 2)] = a * symmetric_uniform(D.seed, pid, 7);
            }
            if (D.amp3 != 0.0) { // single mode on top of the noise (a type-3 perturbation)
                const double v = D.amp3 * sin(D.k3 * z);
                f[herm_re<NF>(0, 1)] += v;
                fb[herm_re<NF>(0, 1)] += v;
            }
            const double ct = D.st4_fh[0] * u[0] + D.st4_fh[1] * u[1] + D.st4_fh[2] * u[2];
            const double ctb = D.st4_fhb[0] * u[0] + D.st4_fhb[1] * u[1] + D.st4_fhb[2] * u[2];
            r[I::N] = D.ndens * D.scale_fac * (1. + 3. * D.st4_ff * ct);
            r[I::Nbar] = D.ndensbar * D.scale_fac * (1. + 3. * D.st4_ffb * ctb);
        } else { // 5: Minerbo closure (:574-645)
            double Nf[3];
            for (int s = 0; s < 3; s++) {
                const double mu = D.ff[s] > 0 ? (D.fdir[s][0] * u[0] + D.fdir[s][1] * u[1] + D.fdir[s][2] * u[2]) / D.ff[s] : 0.0;
                double af = exp(D.Z[s] * mu);
                if (D.Z[s] / 3.0 > 1e-3) af *= D.Z[s] / sinh(D.Z[s]);
                Nf[s] = D.n5[s] * D.scale_fac * af;
            }
            Nf[2] = Nf[2] / 4.0;
            double N = Nf[0] + Nf[2], Nb = Nf[1] + Nf[2];
            if (NF == 3) { N += Nf[2]; Nb += Nf[2]; }
            r[I::N] = N;
            r[I::Nbar] = Nb;
            f[herm_re<NF>(0, 0)] = Nf[0] / N;
            f[herm_re<NF>(1, 1)] = Nf[2] / N;
            fb[herm_re<NF>(0, 0)] = Nf[1] / Nb;
            fb[herm_re<NF>(1, 1)] = Nf[2] / Nb;
            if (NF == 3) {
                f[herm_re<NF>(2, 2)] = Nf[2] / N;
                fb[herm_re<NF>(2, 2)] = Nf[2] / Nb;
            }
            int d = 0;
            auto pert = [&](double* t, int i, int j, int a_, int b_) {
                t[herm_re<NF>(i, j)] = D.amp * symmetric_uniform(D.seed, pid, d++) * (t[herm_re<NF>(a_, a_)] - t[herm_re<NF>(b_, b_)]);
                t[herm_im<NF>(i, j)] = D.amp * symmetric_uniform(D.seed, pid, d++) * (t[herm_re<NF>(a_, a_)] - t[herm_re<NF>(b_, b_)]);
            };
            pert(f, 0, 1, 0, 1);
            pert(fb, 0, 1, 0, 1);
            if (NF == 3) {
                pert(f, 0, 2, 0, 2);
                pert(f, 1, 2, 1, 2);
                pert(fb, 0, 2, 0, 2);
                pert(fb, 1, 2, 1, 2);
            }
        }
        r[I::pupx] = u[0] * r[I::pupt];
        r[I::pupy] = u[1] * r[I::pupt];
        r[I::pupz] = u[2] * r[I::pupt];
        // set_trace_length (generate_code.py:450-457)
        r[I::L] = sqrt(su_mag2<NF, double>(f));
        r[I::Lbar] = sqrt(su_mag2<NF, double>(fb));
        P.pos[0][pid] = x;
        P.pos[1][pid] = y;
        P.pos[2][pid] = z;
        for (int k = 0; k < I::nattribs; k++) P.r[k][pid] = r[k];
        if (ids) ids[pid] = pid + 1; // ParticleType::NextID() starts at 1
    }
}

double minerbo_Z(double fluxfac)
{
    // FlavoredNeutrinoContainerInit.cpp:52-80
    const double maxresidual = 1e-6, minfluxfac = 1e-3;
    const int maxcount = 20;
    double Z = 1.0;
    if (fluxfac < minfluxfac) return 3. * fluxfac;
    double residual = 1.0;
    int count = 0;
    while (std::fabs(residual) > maxresidual && count < maxcount) {
        residual = fluxfac - 1.0 / std::tanh(Z) + 1.0 / Z;
        const double slope = 1.0 / (std::sinh(Z) * std::sinh(Z)) - 1.0 / (Z * Z);
        Z -= residual / slope;
        count++;
    }
    return residual > maxresidual ? NAN : Z;
}

} // namespace

// uniform_sphere_xyz (FlavoredNeutrinoContainerInit.cpp:9-37) and the other sets.
std::vector<double> direction_set(const emu_init_config* c)
{
    std::vector<double> xyz;
    if (c->direction_set == EMU_DIRS_SPHERE) {
        const int n = c->nphi_equator;
        const double dtheta = M_PI * std::sqrt(3) / n;
        double theta = 0, phi0 = 0;
        while (theta < M_PI / 2.) {
            const int nphi = theta == 0 ? n : (int)std::lround(n * std::cos(theta));
            const double dphi = 2. * M_PI / nphi;
            if (nphi == 1) theta = M_PI / 2.;
            for (int iphi = 0; iphi < nphi; iphi++) {
                const double phi = phi0 + iphi * dphi;
                const double x = std::cos(theta) * std::cos(phi), y = std::cos(theta) * std::sin(phi), z = std::sin(theta);
                xyz.insert(xyz.end(), {x, y, z});
                if (theta > 0) xyz.insert(xyz.end(), {-x, -y, -z});
            }
            theta += dtheta;
            phi0 = phi0 + 0.5 * dphi;
        }
    } else if (c->direction_set == EMU_DIRS_CUBE26) {
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    if (!dx && !dy && !dz) continue;
                    const double n = std::sqrt((double)(dx * dx + dy * dy + dz * dz));
                    xyz.insert(xyz.end(), {dx / n, dy / n, dz / n});
                }
    } else {
        const int n = c->ndirections;
        const double ga = M_PI * (3.0 - std::sqrt(5.0));
        for (int i = 0; i < n; i++) {
            const double z = 1.0 - (2.0 * i + 1.0) / n, r = std::sqrt(std::max(0.0, 1.0 - z * z));
            xyz.insert(xyz.end(), {r * std::cos(ga * i), r * std::sin(ga * i), z});
        }
    }
    return xyz;
}

} // namespace emu

using namespace emu;

extern "C" int64_t emu_init_count(const emu_params* p, const emu_init_config* c)
{
    if (!p || !c) return -1;
    const int64_t ndirs = (int64_t)direction_set(c).size() / 3;
    return (int64_t)p->ncell[0] * p->ncell[1] * p->ncell[2] * c->nppc[0] * c->nppc[1] * c->nppc[2] * ndirs;
}

extern "C" int emu_init_particles(emu_ctx* ctx, const emu_init_config* c, emu_particles* out, int64_t* ids)
{
    if (!ctx || !c || !out) return EMU_ERR_ARG;
    const emu_params& p = ctx->p;
    if (c->simulation_type < 0 || c->simulation_type > 5) return set_error(ctx, EMU_ERR_ARG, "Invalid simulation type");
    const std::vector<double> dirs = direction_set(c);
    const int ndirs = (int)(dirs.size() / 3);
    if (emu_init_count(&p, c) != out->np) return set_error(ctx, EMU_ERR_ARG, "particle count mismatch");
    InitDev D{};
    D.nf = p.nf;
    D.st = c->simulation_type;
    D.nx = p.ncell[0];
    D.ny = p.ncell[1];
    D.nz = p.ncell[2];
    for (int d = 0; d < 3; d++) {
        D.nppc[d] = c->nppc[d];
        D.plo[d] = p.plo[d];
        D.dx[d] = ctx->g.dx[d];
    }
    D.ndirs = ndirs;
    D.nlocs = c->nppc[0] * c->nppc[1] * c->nppc[2];
    D.scale_fac = D.dx[0] * D.dx[1] * D.dx[2] / D.nlocs / ndirs; // :156
    D.seed = c->seed;
    const double dm2 = (p.mass[1] - p.mass[0]) * (p.mass[1] - p.mass[0]);
    const double E50 = 50. * 1e6 * phys::eV;
    D.pupt = E50;
    const double mu_ndens = std::sqrt(2.) * phys::GF;
    switch (D.st) {
    case 0: D.pupt = dm2 * phys::c4 * std::sin(2. * p.theta12) / (8. * M_PI * phys::hbarc); break;
    case 1: D.ndens = 10. * dm2 * phys::c4 / (2. * std::sqrt(2.) * phys::GF * E50); break;
    case 2: D.ndens = (dm2 * phys::c4 / (2. * E50)) / (2. * mu_ndens); break;
    case 3: {
        const double lambda = (p.phi[2] - p.plo[2]) / (double)c->st3_wavelength_fraction_of_domain;
        D.k3 = (2. * M_PI) / lambda;
        D.amp = c->st3_amplitude;
        D.ndens = (dm2 * phys::c4 / (2. * E50) + D.k3 * phys::hbarc) / (2. * mu_ndens);
        break;
    }
    case 4:
        D.ndens = c->st4_ndens;
        D.ndensbar = c->st4_ndensbar;
        D.amp = c->st4_amplitude;
        D.st4_fh[0] = std::cos(c->st4_phi) * std::sin(c->st4_theta);
        D.st4_fh[1] = std::sin(c->st4_phi) * std::sin(c->st4_theta);
        D.st4_fh[2] = std::cos(c->st4_theta);
        D.st4_fhb[0] = std::cos(c->st4_phibar) * std::sin(c->st4_thetabar);
        D.st4_fhb[1] = std::sin(c->st4_phibar) * std::sin(c->st4_thetabar);
        D.st4_fhb[2] = std::cos(c->st4_thetabar);
        D.st4_ff = c->st4_fluxfac;
        D.st4_ffb = c->st4_fluxfacbar;
        D.amp3 = c->st3_amplitude;
        if (D.amp3 != 0.0)
            D.k3 = (2. * M_PI) / ((p.phi[2] - p.plo[2]) / (double)c->st3_wavelength_fraction_of_domain);
        break;
    case 5: {
        const double n5[3] = {c->st5[0], c->st5[1], c->st5[2]};
        for (int s = 0; s < 3; s++) {
            D.n5[s] = n5[s];
            D.fdir[s][0] = c->st5[3 + s];
            D.fdir[s][1] = c->st5[6 + s];
            D.fdir[s][2] = c->st5[9 + s];
            D.ff[s] = std::sqrt(D.fdir[s][0] * D.fdir[s][0] + D.fdir[s][1] * D.fdir[s][1] + D.fdir[s][2] * D.fdir[s][2]);
            D.Z[s] = minerbo_Z(D.ff[s]);
            if (std::isnan(D.Z[s])) return set_error(ctx, EMU_ERR_ARG, "Failed to converge on a solution.");
        }
        D.amp = c->st5[12];
        break;
    }
    }
    double* ddirs = nullptr;
    EMU_HIP(ctx, hipMalloc(&ddirs, std::max<size_t>(dirs.size(), 3) * sizeof(double)));
    if (!dirs.empty()) EMU_HIP(ctx, hipMemcpyAsync(ddirs, dirs.data(), dirs.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    D.dirs = ddirs;
    PView P;
    P.np = out->np;
    const int na = 12 + 2 * p.nf * p.nf;
    for (int d = 0; d < 3; d++) P.pos[d] = out->pos[d];
    for (int k = 0; k < EMU_MAX_ATTRIBS; k++) P.r[k] = k < na ? out->rdata[k] : nullptr;
    if (P.np > 0) {
        const unsigned nb = (unsigned)std::min<int64_t>(65536, (P.np + 255) / 256);
        if (p.nf == 2) hipLaunchKernelGGL(k_init<2>, dim3(nb), dim3(256), 0, ctx->stream, D, P, ids);
        else hipLaunchKernelGGL(k_init<3>, dim3(nb), dim3(256), 0, ctx->stream, D, P, ids);
    }
    const hipError_t e = hipGetLastError();
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ddirs);
    if (e != hipSuccess) return set_hip_error(ctx, e, "init_particles");
    return EMU_OK;
}
