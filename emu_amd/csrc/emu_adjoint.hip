// emu_adjoint.hip -- the flavour right-hand side and its vector-Jacobian
// product (VJP), for reverse-mode gradients through the RK4 evolution.
//
// Particles stay in a fixed order (gradients are per particle).  The gathers
// take one particle per thread with its own stencil; the deposits sum over
// the particles that share a stencil on the matrix cores first
// (k_cell_deposit), then add to the grid with FP64 atomics.  All arrays are
// SoA: [component][np].
//
// Forward RHS for stage state Y (positions frozen within the step):
//   q_p  = N_p f_p - conj(Nbar_p fbar_p)                    (per particle)
//   M_t  = base + sum_p w_t(p) q_p,  Phi_t = sum_p w_t(p) v_p q_p      (deposit)
//   s_p  = coef sum_t w_t(p) (M_t - v_p . Phi_t)                       (gather)
//   k_p  = -i [Hvac_p + s_p, f_p],  kbar_p = -i [conj(Hvac_p - s_p), fbar_p]
// VJP, given lambda (the adjoint of k):
//   the commutator K = -i[A, B] with L = lambda . K = tr(K Lam) (Lam the
//   Hermitian matrix of lambda with halved off-diagonals) has
//   dL/dB = -i[Lam, A] and dL/dA = -i[B, Lam] as matrices; as components the
//   off-diagonal entries count twice.
//   mu_p = coef (dL/dH_p + sigma dL/dHbar_p), sigma = -1 (Re), +1 (Im)
//   scatter (adjoint of the gather): LM_t += w_t(p) mu_p, LPhi_t -= w_t(p) v_p mu_p
//   gather (adjoint of the deposit): Lq_p = sum_t w_t(p) (LM_t + v_p . LPhi_t)
//   dL/df_p += N_p Lq_p;  dL/dfbar_p += sigma Nbar_p Lq_p;  dL/dNbar_p += sigma fbar_p . Lq_p
#include "emu_internal.h"
#include "emu_stage.h"
#include "../../include/emu_adjoint.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

namespace emu {
namespace {

constexpr int kB = 256;

unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kB - 1) / kB, 1 << 20)); }

// f(cell, weight) over a particle's stencil in the reference's order (k outer,
// then j, then i; Evolve.cpp:118-124): 3 points per dimension where the shape
// order is above 0, else the particle's own cell.  Unrolled, with the wrapped
// indices of each dimension computed once, so nothing is indexed at run time.
template <typename F> __device__ __forceinline__ void for_stencil(const Geom& g, const double pos[3], F&& f)
{
    int idx[3], cw[3][3];
    double wd[3][3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        shape3((pos[d] - g.plo[d]) * g.dxi[d], g.order[d], idx[d], wd[d]);
#pragma unroll
        for (int o = 0; o < 3; o++) cw[d][o] = pmod(idx[d] + o - 1, g.n[d]);
    }
    const bool w3[3] = {g.order[0] > 0, g.order[1] > 0, g.order[2] > 0};
#pragma unroll
    for (int kk = 0; kk < 3; kk++) {
        if (!w3[2] && kk != 1) continue;
#pragma unroll
        for (int jj = 0; jj < 3; jj++) {
            if (!w3[1] && jj != 1) continue;
#pragma unroll
            for (int ii = 0; ii < 3; ii++) {
                if (!w3[0] && ii != 1) continue;
                f(((int64_t)cw[2][kk] * g.n[1] + cw[1][jj]) * g.n[0] + cw[0][ii], wd[0][ii] * wd[1][jj] * wd[2][kk]);
            }
        }
    }
}

// Particle inputs; T is the storage of the flavour arrays and particle
// constants (double, or float for the FP32-bulk step calls), positions stay
// double.
template <typename T> struct AdjInT {
    int64_t np;
    const double* pos; // [3][np]
    const T* pup;      // [4][np]
    const T* Nn;       // [2][np]
    const T* Y;        // [2 NH][np]
};
using AdjIn = AdjInT<double>;

template <int NF>
__global__ void k_fill_base(double* grid, int64_t ncell, double m00)
{
    const int64_t n = ncell * 4 * NH<NF>;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        grid[t] = (t % (4 * NH<NF>)) == 0 ? m00 : 0.0;
}

struct HamParams {
    double vac[9];
    double coef;
};

// H/hbar and Hbar/hbar of a particle from the deposited grid
template <int NF>
__device__ __forceinline__ void hamiltonians(const AdjIn& in, int64_t p, const Geom& g, const double* grid,
                                             const HamParams& hp, double v[3], double* H, double* Hb)
{
    constexpr int nh = NH<NF>, QG = 4 * nh;
    const int64_t np = in.np;
    const double pos[3] = {in.pos[p], in.pos[np + p], in.pos[2 * np + p]};
    const double pt = in.pup[3 * np + p];
    v[0] = in.pup[p] / pt;
    v[1] = in.pup[np + p] / pt;
    v[2] = in.pup[2 * np + p] / pt;
    double inside[nh];
#pragma unroll
    for (int c = 0; c < nh; c++) inside[c] = 0.0;
    for_stencil(g, pos, [&](int64_t cell, double w) {
        const double* G = grid + cell * QG;
#pragma unroll
        for (int c = 0; c < nh; c++)
            inside[c] += w * (G[c] - v[0] * G[nh + c] - v[1] * G[2 * nh + c] - v[2] * G[3 * nh + c]);
    });
    for (int c = 0; c < nh; c++) {
        const double hv = hp.vac[c] / pt, hs = hp.coef * inside[c];
        H[c] = hv + hs;
        Hb[c] = is_im_comp<NF>(c) ? hs - hv : hv - hs;
    }
}

// (perm, when given, takes the particles in cell order: the lanes of a wave then
// read the same grid cells)
template <int NF>
__global__ __launch_bounds__(256) void k_rhs(AdjIn in, Geom g, const int* perm, const double* grid, HamParams hp, double* k)
{
    constexpr int nh = NH<NF>;
    const int64_t np = in.np;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = perm ? perm[i] : i;
        double v[3], H[nh], Hb[nh], f[nh], out[nh];
        hamiltonians<NF>(in, p, g, grid, hp, v, H, Hb);
        for (int c = 0; c < nh; c++) f[c] = in.Y[(int64_t)c * np + p];
        commutator_rhs<NF, double>(H, f, out, 1.0);
        for (int c = 0; c < nh; c++) k[(int64_t)c * np + p] = out[c];
        for (int c = 0; c < nh; c++) f[c] = in.Y[(int64_t)(nh + c) * np + p];
        commutator_rhs<NF, double>(Hb, f, out, 1.0);
        for (int c = 0; c < nh; c++) k[(int64_t)(nh + c) * np + p] = out[c];
    }
}

// Hermitian component vector of Lam: off-diagonals halved.
template <int NF, typename T> __device__ __forceinline__ void lam_matrix(const T* lam, T* L)
{
    for (int c = 0; c < NH<NF>; c++) L[c] = lam[c];
    for (int a = 0; a < NF; a++)
        for (int b = a + 1; b < NF; b++) {
            L[herm_re<NF>(a, b)] *= T(0.5);
            L[herm_im<NF>(a, b)] *= T(0.5);
        }
}

template <int NF, typename T> __device__ __forceinline__ void double_offdiag(T* x)
{
    for (int a = 0; a < NF; a++)
        for (int b = a + 1; b < NF; b++) {
            x[herm_re<NF>(a, b)] *= T(2);
            x[herm_im<NF>(a, b)] *= T(2);
        }
}

// VJP part 1: commutator adjoints.  lamY <- dL/dY through [H, .];
// mu <- the adjoint of the gathered field (per component).
template <int NF>
__global__ __launch_bounds__(256) void k_vjp_comm(AdjIn in, Geom g, const int* perm, const double* grid, HamParams hp, const double* lamk,
                           double* lamY, double* mu)
{
    constexpr int nh = NH<NF>;
    const int64_t np = in.np;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = perm ? perm[i] : i;
        double v[3], H[nh], Hb[nh];
        hamiltonians<NF>(in, p, g, grid, hp, v, H, Hb);
        double gH[2][nh];
        for (int bar = 0; bar < 2; bar++) {
            double lam[nh], L[nh], f[nh], gy[nh];
            for (int c = 0; c < nh; c++) {
                lam[c] = lamk[(int64_t)(bar * nh + c) * np + p];
                f[c] = in.Y[(int64_t)(bar * nh + c) * np + p];
            }
            lam_matrix<NF>(lam, L);
            commutator_rhs<NF, double>(L, bar ? Hb : H, gy, 1.0); // -i[Lam, A]
            double_offdiag<NF>(gy);
            for (int c = 0; c < nh; c++) lamY[(int64_t)(bar * nh + c) * np + p] = gy[c];
            commutator_rhs<NF, double>(f, L, gH[bar], 1.0);       // -i[B, Lam]
            double_offdiag<NF>(gH[bar]);
        }
        for (int c = 0; c < nh; c++)
            mu[(int64_t)c * np + p] = hp.coef * (gH[0][c] + (is_im_comp<NF>(c) ? gH[1][c] : -gH[1][c]));
    }
}

// The centre cell of each particle's stencil, and the particle count per cell.
// The lanes of a wave holding one key add to its counter once (particles in
// cell order give one or two keys per wave: per-lane atomics on one counter
// would serialise).
template <typename F> __device__ __forceinline__ void for_each_wave_key(bool act, int key, F&& f)
{
    unsigned long long left = __ballot(act);
    while (left) {
        const int lead = __ffsll((long long)left) - 1;
        const int k = __shfl(key, lead);
        const unsigned long long m = __ballot(act && key == k);
        f(k, m, lead);
        left &= ~m;
    }
}

__global__ void k_cell_keys(int64_t np, const double* pos, Geom g, int* key, int* counts)
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool act = p < np;
    const int lane = threadIdx.x & 63;
    int k = 0;
    if (act) {
        int c[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            int idx;
            double w[3];
            shape3((pos[d * np + p] - g.plo[d]) * g.dxi[d], g.order[d], idx, w);
            c[d] = pmod(idx, g.n[d]);
        }
        k = (c[2] * g.n[1] + c[1]) * g.n[0] + c[0];
        key[p] = k;
    }
    for_each_wave_key(act, k, [&](int kk, unsigned long long m, int lead) {
        if (lane == lead) atomicAdd(&counts[kk], __popcll(m));
    });
}

// perm[start[key] + i] = p: the particles listed cell by cell; within a cell in
// lane order within a wave, waves in the order their atomics land (which does
// not matter to the deposit's sums beyond rounding).
__global__ void k_cell_perm(int64_t np, const int* key, const int* start, int* cursor, int* perm)
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool act = p < np;
    const int lane = threadIdx.x & 63;
    const int k = act ? key[p] : 0;
    for_each_wave_key(act, k, [&](int kk, unsigned long long m, int lead) {
        int base = 0;
        if (lane == lead) base = atomicAdd(&cursor[kk], __popcll(m));
        base = __shfl(base, lead);
        if ((m >> lane) & 1ull) perm[start[kk] + base + __popcll(m & ((1ull << lane) - 1ull))] = (int)p;
    });
}

// VJP part 3: gather of the grid adjoint (adjoint of the deposit).
template <int NF>
__global__ __launch_bounds__(256) void k_vjp_gather(AdjIn in, Geom g, const int* perm, const double* gadj, double* lamY, double* lamNbar)
{
    constexpr int nh = NH<NF>, QG = 4 * nh;
    const int64_t np = in.np;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = perm ? perm[i] : i;
        const double pos[3] = {in.pos[p], in.pos[np + p], in.pos[2 * np + p]};
        const double pt = in.pup[3 * np + p];
        const double v[3] = {in.pup[p] / pt, in.pup[np + p] / pt, in.pup[2 * np + p] / pt};
        double lq[nh];
#pragma unroll
        for (int c = 0; c < nh; c++) lq[c] = 0.0;
        for_stencil(g, pos, [&](int64_t cell, double w) {
            const double* G = gadj + cell * QG;
#pragma unroll
            for (int c = 0; c < nh; c++)
                lq[c] += w * (G[c] + v[0] * G[nh + c] + v[1] * G[2 * nh + c] + v[2] * G[3 * nh + c]);
        });
        const double N = in.Nn[p], Nb = in.Nn[np + p];
        double ln = 0.0;
        for (int c = 0; c < nh; c++) {
            const double sg = is_im_comp<NF>(c) ? 1.0 : -1.0;
            lamY[(int64_t)c * np + p] += N * lq[c];
            lamY[(int64_t)(nh + c) * np + p] += sg * Nb * lq[c];
            ln += sg * in.Y[(int64_t)(nh + c) * np + p] * lq[c];
        }
        if (lamNbar) lamNbar[p] += ln;
    }
}

// Objective J = sum over cells of |N_01|^2 with N_01 = sum_p w N f01 (nu).
// One thread per particle; the lanes of a wave whose stencils share a centre
// cell (one or two such groups per wave in cell order) sum each target's
// contribution across the wave before one atomic per target and component:
// per-particle atomics would put every particle of a cell on the same 27
// targets.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

template <int NF, typename T>
__global__ __launch_bounds__(256) void k_obj_deposit(AdjInT<T> in, Geom g, double* g2)
{
    const int64_t np = in.np;
    const int lane = threadIdx.x & 63;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    // p0 is the wave's first particle, so the loop bound is the same for all its lanes
    for (int64_t p0 = blockIdx.x * (int64_t)blockDim.x + (threadIdx.x & ~63); p0 < np; p0 += stride) {
        const int64_t p = p0 + lane;
        const bool act = p < np;
        int c[3] = {0, 0, 0};
        double wd[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
        double re = 0.0, im = 0.0;
        if (act) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                int idx;
                shape3((in.pos[d * np + p] - g.plo[d]) * g.dxi[d], g.order[d], idx, wd[d]);
                c[d] = pmod(idx, g.n[d]);
            }
            const double N = in.Nn[p];
            re = N * double(in.Y[(int64_t)herm_re<NF>(0, 1) * np + p]);
            im = N * double(in.Y[(int64_t)herm_im<NF>(0, 1) * np + p]);
        }
        const int key = (c[2] * g.n[1] + c[1]) * g.n[0] + c[0];
        for_each_wave_key(act, key, [&](int k, unsigned long long m, int lead) {
            const bool mine = (m >> lane) & 1ull;
            const int kx = k % g.n[0], ky = (k / g.n[0]) % g.n[1], kz = k / (g.n[0] * g.n[1]);
#pragma unroll
            for (int kk = 0; kk < 3; kk++) {
                if (g.order[2] == 0 && kk != 1) continue;
                const int tz = pmod(kz + kk - 1, g.n[2]);
#pragma unroll
                for (int jj = 0; jj < 3; jj++) {
                    if (g.order[1] == 0 && jj != 1) continue;
                    const int ty = pmod(ky + jj - 1, g.n[1]);
#pragma unroll
                    for (int ii = 0; ii < 3; ii++) {
                        if (g.order[0] == 0 && ii != 1) continue;
                        const int64_t t = ((int64_t)tz * g.n[1] + ty) * g.n[0] + pmod(kx + ii - 1, g.n[0]);
                        const double w = wd[0][ii] * wd[1][jj] * wd[2][kk];
                        const double a = wave_sum(mine ? w * re : 0.0), b = wave_sum(mine ? w * im : 0.0);
                        if (lane == lead) {
                            unsafeAtomicAdd(&g2[2 * t], a);
                            unsafeAtomicAdd(&g2[2 * t + 1], b);
                        }
                    }
                }
            }
        });
    }
}

__global__ void k_obj_sum(const double* g2, int64_t ncell, double* J)
{
    double s = 0.0;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x)
        s += g2[2 * c] * g2[2 * c] + g2[2 * c + 1] * g2[2 * c + 1];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) unsafeAtomicAdd(J, s);
}

template <int NF, typename T>
__global__ void k_obj_grad(AdjInT<T> in, Geom g, const double* g2, T* dJdY)
{
    const int64_t np = in.np;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x) {
        const double pos[3] = {in.pos[p], in.pos[np + p], in.pos[2 * np + p]};
        double re = 0.0, im = 0.0;
        for_stencil(g, pos, [&](int64_t cell, double w) {
            re += w * g2[2 * cell];
            im += w * g2[2 * cell + 1];
        });
        const double N = in.Nn[p];
        for (int c = 0; c < 2 * NH<NF>; c++) dJdY[(int64_t)c * np + p] = T(0);
        dJdY[(int64_t)herm_re<NF>(0, 1) * np + p] = T(2.0 * N * re);
        dJdY[(int64_t)herm_im<NF>(0, 1) * np + p] = T(2.0 * N * im);
    }
}

__global__ void k_positions(int64_t np, double* pos, const double* pup, double dt, Geom g)
{
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x) {
        double x[3], v[3], nx[3];
        for (int d = 0; d < 3; d++) {
            x[d] = pos[(int64_t)d * np + p];
            v[d] = pup[(int64_t)d * np + p] / pup[3 * np + p] * phys::c;
        }
        int nc;
        end_position(x, v, dt, g, nx, nc);
        for (int d = 0; d < 3; d++) pos[(int64_t)d * np + p] = nx[d];
    }
}

template <int NF>
__global__ void k_renorm_soa(int64_t np, double* Y, const double* Ll, double maxError, unsigned long long* err)
{
    constexpr int nh = NH<NF>;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < np; p += (int64_t)gridDim.x * blockDim.x) {
        for (int bar = 0; bar < 2; bar++) {
            double f[nh];
            for (int c = 0; c < nh; c++) f[c] = Y[(int64_t)(bar * nh + c) * np + p];
            const int e = renorm_fp64<NF>(f, Ll[(int64_t)bar * np + p], maxError);
            if (e != EMU_OK) {
                atomicMin(err, status_word(p, e));
                break;
            }
            for (int c = 0; c < nh; c++) Y[(int64_t)(bar * nh + c) * np + p] = f[c];
        }
    }
}

// ---------------------------------------------------------------------------
// Step-level kernels: one launch per RK4 stage after that stage's deposit.
// The stage algebra is step()'s (AMReX RK4 with the reference's saxpys,
// FlavoredNeutrinoContainer.H:39-46), each a x + y rounded as written.
//
// They take one cell per wave (the cell list's runs): every particle of the
// wave has its stencil in the cell's 27-cell neighbourhood, so the grid rows
// are the same for all lanes and are read as scalar loads, one neighbour at
// a time, while each lane forms its own weights.
// ---------------------------------------------------------------------------
constexpr int kCellWaves = 4; // waves (cells) per workgroup

// The wave's cell, in an XCD-aware order (consecutive cells share an XCD's
// L2); -1 past the last cell.
__device__ __forceinline__ int wave_cell(int64_t ncell, int blocks_per_xcd)
{
    const int b = blockIdx.x;
    const int64_t c = ((int64_t)(b & 7) * blocks_per_xcd + (b >> 3)) * kCellWaves + (threadIdx.x >> 6);
    return c < ncell ? __builtin_amdgcn_readfirstlane((int)c) : -1;
}

// The deposit of both directions with the sum over particles on the matrix
// cores.  KIND 0: the forward q_c = N f_c - conj(Nbar fbar_c) with (1, v);
// KIND 1: the gather's adjoint mu_c with (1, -v).  Per-particle atomics would
// put every particle of a cell on the same 27 x 4 NH grid entries, so the sum
// over a cell's particles is done first, as in the fused stage kernel.
//
// One wave per cell of the cell list (the centre cell of the particles'
// stencils): its particles share the 27 targets, so D[t][q] = sum_p W[p][t]
// R[p][q] runs on f64 16x16x4 MFMAs over all of them, 64 at a time, and is
// added to the grid once per cell, one atomic per value.
template <int NF, int WX, int WY, int WZ, int KIND, typename T>
__global__ __launch_bounds__(64 * kCellWaves) void k_cell_deposit(AdjInT<T> in, Geom g, const int* start,
                                                                  const int* perm, int64_t ncell, int bpx,
                                                                  const T* mu, double* grid)
{
    constexpr int nh = NH<NF>, QG = 4 * nh;
    constexpr int NT = WX * WY * WZ, MT = (NT + 15) / 16;
    // full 16-column tiles (f64 16x16x4), then the last columns (NF = 3: 4 of
    // 36) as 4-column tiles of the four-block f64 4x4x4 form: 16 cycles in
    // place of a 64-cycle 16x16x4 step over 12 padding columns
    constexpr int NTF = QG / 16, NSM = (QG - 16 * NTF + 3) / 4;
    constexpr int NSR = nh + 4;            // q (or mu), v (or -v), 1
    constexpr int REG = 64 * 9 + 64 * NSR; // doubles per wave: weights, sources
    __shared__ double lds[kCellWaves * REG];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double* Wf = lds + wave * REG;
    double* Sr = Wf + 64 * 9;
    const int cell = wave_cell(ncell, bpx);
    if (cell < 0) return;
    const int beg = __builtin_amdgcn_readfirstlane(start[cell]), end = __builtin_amdgcn_readfirstlane(start[cell + 1]);
    if (beg == end) return;
    const int64_t np = in.np;
    // operand lane constants, as in the stage kernel: row t = 16m + l%16 reads the
    // shape factors of target t; column q = f nh + c is source c times (1, v)_f
    int tdx[MT][3];
#pragma unroll
    for (int m = 0; m < MT; m++) {
        const int t = 16 * m + (lane & 15);
        const int tc = t < NT ? t : NT - 1;
        const int ii = tc % WX, jj = (tc / WX) % WY, kk = tc / (WX * WY);
        tdx[m][0] = WX == 1 ? 1 : ii;
        tdx[m][1] = 3 + (WY == 1 ? 1 : jj);
        tdx[m][2] = 6 + (WZ == 1 ? 1 : kk);
    }
    // B column q (of a 16-column tile: 16n + l%16; of a 4-column one:
    // 16 NTF + 4k + l%4) is source c times multiplier f, q = f nh + c
    constexpr int NCG = NTF + NSM;
    int bi1[NCG], bi2[NCG];
#pragma unroll
    for (int n = 0; n < NCG; n++) {
        const int q = n < NTF ? 16 * n + (lane & 15) : 16 * NTF + 4 * (n - NTF) + (lane & 3);
        const int f = q / nh, c = q % nh;
        bi1[n] = q < QG ? c : nh + 3;
        bi2[n] = q < QG ? (f == 0 ? nh + 3 : nh + f - 1) : nh + 3;
    }
    mf_d4 acc[MT][NTF];
    double accs[MT][NSM > 0 ? NSM : 1];
#pragma unroll
    for (int m = 0; m < MT; m++) {
#pragma unroll
        for (int n = 0; n < NTF; n++) acc[m][n] = mf_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < NSM; k++) accs[m][k] = 0.0;
    }
    for (int base = beg; base < end; base += 64) {
        // stage the chunk: every lane a valid particle, the inactive ones with zero sources
        const bool act = base + lane < end;
        const int64_t pl = perm[act ? base + lane : beg];
        const double pos[3] = {in.pos[pl], in.pos[np + pl], in.pos[2 * np + pl]};
        const T pt = in.pup[3 * np + pl];
        const T v[3] = {in.pup[pl] / pt, in.pup[np + pl] / pt, in.pup[2 * np + pl] / pt};
        double wd[3][3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            int idx;
            shape3((pos[d] - g.plo[d]) * g.dxi[d], g.order[d], idx, wd[d]);
        }
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int e = 0; e < 3; e++) Wf[lane * 9 + 3 * d + e] = wd[d][e];
        // the sources in T (FP32-rounded in the FP32 form, as its fused
        // deposit's), summed in FP64 on the matrix cores
        if constexpr (KIND == 0) {
            const T N = in.Nn[pl], Nb = in.Nn[np + pl];
#pragma unroll
            for (int c = 0; c < nh; c++) {
                const T f = in.Y[(int64_t)c * np + pl], fb = in.Y[(int64_t)(nh + c) * np + pl];
                Sr[lane * NSR + c] = act ? double(is_im_comp<NF>(c) ? N * f + Nb * fb : N * f - Nb * fb) : 0.0;
            }
        } else {
#pragma unroll
            for (int c = 0; c < nh; c++) Sr[lane * NSR + c] = act ? double(mu[(int64_t)c * np + pl]) : 0.0;
        }
#pragma unroll
        for (int d = 0; d < 3; d++) Sr[lane * NSR + nh + d] = double(KIND == 0 ? v[d] : -v[d]);
        Sr[lane * NSR + nh + 3] = 1.0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // k-steps of 4 particles, two per iteration (the second at immediate
        // offsets from the same per-lane pointers); an odd count runs one more
        // over rows with zero sources, which add nothing
        int nks = (min(64, end - base) + 3) >> 2;
        nks += nks & 1;
        const double* wl = Wf + (lane >> 4) * 9;
        const double* sl = Sr + (lane >> 4) * NSR;
        for (int ks = 0; ks < nks; ks += 2) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const double* wf = wl + (ks + h) * 4 * 9;
                const double* sr = sl + (ks + h) * 4 * NSR;
                double A[MT];
#pragma unroll
                for (int m = 0; m < MT; m++) A[m] = wf[tdx[m][0]] * wf[tdx[m][1]] * wf[tdx[m][2]];
#pragma unroll
                for (int n = 0; n < NCG; n++) {
                    const double B = sr[bi1[n]] * sr[bi2[n]];
                    if (n < NTF) {
#pragma unroll
                        for (int m = 0; m < MT; m++)
                            acc[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[m], B, acc[m][n], 0, 0, 0);
                    } else {
#pragma unroll
                        for (int m = 0; m < MT; m++)
                            accs[m][n - NTF] = __builtin_amdgcn_mfma_f64_4x4x4f64(A[m], B, accs[m][n - NTF], 0, 0, 0);
                    }
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // element r of lane l holds D[16m + 4r + l/16][16n + l%16]; the neighbours
    // of a cell are at most one period away, so they wrap with one add or
    // subtract
    const int kx = __builtin_amdgcn_readfirstlane(cell % g.n[0]);
    const int ky = __builtin_amdgcn_readfirstlane((cell / g.n[0]) % g.n[1]);
    const int kz = __builtin_amdgcn_readfirstlane(cell / (g.n[0] * g.n[1]));
    auto wrap = [](int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); };
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int t = 16 * m + 4 * r + (lane >> 4);
            if (t >= NT) continue;
            const int ii = t % WX, jj = (t / WX) % WY, kk = t / (WX * WY);
            const int tx = WX == 1 ? kx : wrap(kx + ii - 1, g.n[0]);
            const int ty = WY == 1 ? ky : wrap(ky + jj - 1, g.n[1]);
            const int tz = WZ == 1 ? kz : wrap(kz + kk - 1, g.n[2]);
            double* G = grid + (((int64_t)tz * g.n[1] + ty) * g.n[0] + tx) * QG;
#pragma unroll
            for (int n = 0; n < NTF; n++) unsafeAtomicAdd(&G[16 * n + (lane & 15)], acc[m][n][r]);
        }
    // 4x4x4 tiles: element of lane l is D[16m + 4((l >> 2) & 3) + l/16][16 NTF + 4k + l%4]
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int k = 0; k < NSM; k++) {
            const int t = 16 * m + 4 * ((lane >> 2) & 3) + (lane >> 4), q = 16 * NTF + 4 * k + (lane & 3);
            if (t >= NT || q >= QG) continue;
            const int ii = t % WX, jj = (t / WX) % WY, kk = t / (WX * WY);
            const int tx = WX == 1 ? kx : wrap(kx + ii - 1, g.n[0]);
            const int ty = WY == 1 ? ky : wrap(ky + jj - 1, g.n[1]);
            const int tz = WZ == 1 ? kz : wrap(kz + kk - 1, g.n[2]);
            unsafeAtomicAdd(&grid[(((int64_t)tz * g.n[1] + ty) * g.n[0] + tx) * QG + q], accs[m][k]);
        }
}


// out[c] += sum over the neighbourhood of `cell` (wave-uniform) of
// w_t (G_t[0][c] + s0 G_t[1][c] + s1 G_t[2][c] + s2 G_t[3][c]), with the
// particle's shape factors w (shape3's windows) and multipliers s.
template <int NF, typename T>
__device__ __forceinline__ void gather_cell_box(const Geom& g, int cell, const T* grid_, const T w[3][3], const T s[3],
                                                T* out)
{
    constexpr int nh = NH<NF>, QG = 4 * nh;
    typedef __attribute__((address_space(4))) const T* cdp;
    const cdp grid = (cdp)grid_;
    // the neighbourhood's indices are wave-uniform: kept in scalar registers so
    // that the grid rows below are scalar loads (vector loads of one address
    // would return 64 copies of each row through the vector memory path)
    const int cx = __builtin_amdgcn_readfirstlane(cell % g.n[0]);
    const int cy = __builtin_amdgcn_readfirstlane((cell / g.n[0]) % g.n[1]);
    const int cz = __builtin_amdgcn_readfirstlane(cell / (g.n[0] * g.n[1]));
    auto wrap = [](int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); };
    int lo[3], hi[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        lo[d] = g.order[d] > 0 ? 0 : 1;
        hi[d] = g.order[d] > 0 ? 3 : 2;
    }
    auto pick = [](const T* x, int i) { return i == 0 ? x[0] : (i == 1 ? x[1] : x[2]); };
#pragma unroll 1
    for (int kk = lo[2]; kk < hi[2]; kk++) {
        const int nz = wrap(cz + kk - 1, g.n[2]);
#pragma unroll 1
        for (int jj = lo[1]; jj < hi[1]; jj++) {
            const int ny = wrap(cy + jj - 1, g.n[1]);
            const T wyz = pick(w[1], jj) * pick(w[2], kk);
#pragma unroll 1
            for (int ii = lo[0]; ii < hi[0]; ii++) {
                const int64_t nb = __builtin_amdgcn_readfirstlane((nz * g.n[1] + ny) * g.n[0] + wrap(cx + ii - 1, g.n[0]));
                const T wt = pick(w[0], ii) * wyz;
                const T a[4] = {wt, wt * s[0], wt * s[1], wt * s[2]};
                const cdp G = grid + nb * QG;
#pragma unroll
                for (int v = 0; v < 4; v++)
#pragma unroll
                    for (int c = 0; c < nh; c++) out[c] += a[v] * G[v * nh + c];
            }
        }
    }
}

// The particle's shape factors (from the FP64 positions, rounded to T) and
// direction (positions frozen within a step).
template <typename T>
__device__ __forceinline__ void particle_shape(const AdjInT<T>& in, int64_t p, const Geom& g, T w[3][3], T v[3])
{
    const int64_t np = in.np;
    const double x[3] = {in.pos[p], in.pos[np + p], in.pos[2 * np + p]};
    const T pt = in.pup[3 * np + p];
    v[0] = in.pup[p] / pt;
    v[1] = in.pup[np + p] / pt;
    v[2] = in.pup[2 * np + p] / pt;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        int idx;
        double wd[3];
        shape3((x[d] - g.plo[d]) * g.dxi[d], g.order[d], idx, wd);
#pragma unroll
        for (int e = 0; e < 3; e++) w[d][e] = T(wd[e]);
    }
}

// H/hbar and Hbar/hbar of a particle of the wave's cell (as hamiltonians()).
template <int NF, typename T>
__device__ __forceinline__ void cell_hamiltonians(const AdjInT<T>& in, int64_t p, const Geom& g, int cell,
                                                  const T* grid, const HamParams& hp, T* H, T* Hb)
{
    constexpr int nh = NH<NF>;
    T w[3][3], v[3];
    particle_shape(in, p, g, w, v);
    const T s[3] = {-v[0], -v[1], -v[2]};
    T inside[nh];
#pragma unroll
    for (int c = 0; c < nh; c++) inside[c] = T(0);
    gather_cell_box<NF, T>(g, cell, grid, w, s, inside);
    const T pt = in.pup[3 * in.np + p];
#pragma unroll
    for (int c = 0; c < nh; c++) {
        const T hv = T(hp.vac[c]) / pt, hs = T(hp.coef) * inside[c];
        H[c] = hv + hs;
        Hb[c] = is_im_comp<NF>(c) ? hs - hv : hv - hs;
    }
}

// A wave's deposit of its cell's particles on the f64 matrix cores, a chunk
// of up to 64 at a time: lane l's particle adds its source row (NH values,
// zero on inactive lanes) times (1, s_x, s_y, s_z) with its shape factors to
// the cell's 3 x 3 x 3 neighbourhood, D[t][q] = sum_p W[p][t] R[p][q]; flush()
// adds D to the grid with one atomic per value.  An order-0 dimension's
// weights are (0, 1, 0), so its other targets receive zeros.
template <int NF> struct WaveDeposit {
    static constexpr int nh = NH<NF>, QG = 4 * nh, NT = 27, MT = 2;
    // full 16-column tiles, then the last QG % 16 columns as 4-column tiles of
    // the four-block f64 4x4x4 form (as k_cell_deposit)
    static constexpr int NTF = QG / 16, NSM = (QG - 16 * NTF + 3) / 4, NCG = NTF + NSM;
    static constexpr int NSR = nh + 4;            // the source row, s, 1
    static constexpr int REG = 64 * 9 + 64 * NSR; // LDS doubles per wave
    mf_d4 acc[MT][NTF];
    double accs[MT][NSM > 0 ? NSM : 1];
    int tdx[MT][3], bi1[NCG], bi2[NCG];

    __device__ __forceinline__ void init(int lane)
    {
#pragma unroll
        for (int m = 0; m < MT; m++) {
            const int t = 16 * m + (lane & 15);
            const int tc = t < NT ? t : NT - 1;
            tdx[m][0] = tc % 3;
            tdx[m][1] = 3 + (tc / 3) % 3;
            tdx[m][2] = 6 + tc / 9;
        }
#pragma unroll
        for (int n = 0; n < NCG; n++) {
            const int q = n < NTF ? 16 * n + (lane & 15) : 16 * NTF + 4 * (n - NTF) + (lane & 3);
            const int f = q / nh, c = q % nh;
            bi1[n] = q < QG ? c : nh + 3;
            bi2[n] = q < QG ? (f == 0 ? nh + 3 : nh + f - 1) : nh + 3;
        }
#pragma unroll
        for (int m = 0; m < MT; m++) {
#pragma unroll
            for (int n = 0; n < NTF; n++) acc[m][n] = mf_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int k = 0; k < NSM; k++) accs[m][k] = 0.0;
        }
    }

    // every lane of the wave calls this (count: the chunk's particles; the
    // rows of lanes past it must hold zero sources)
    template <typename T>
    __device__ __forceinline__ void add(double* Wf, double* Sr, int lane, const double wd[3][3], const T* src,
                                        const T sv[3], int count)
    {
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int e = 0; e < 3; e++) Wf[lane * 9 + 3 * d + e] = wd[d][e];
#pragma unroll
        for (int c = 0; c < nh; c++) Sr[lane * NSR + c] = double(src[c]);
#pragma unroll
        for (int d = 0; d < 3; d++) Sr[lane * NSR + nh + d] = double(sv[d]);
        Sr[lane * NSR + nh + 3] = 1.0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // two k-steps of 4 particles per iteration, the second at immediate
        // offsets; an odd count runs one more over zero-source rows
        int nks = (min(64, count) + 3) >> 2;
        nks += nks & 1;
        const double* wl = Wf + (lane >> 4) * 9;
        const double* sl = Sr + (lane >> 4) * NSR;
        for (int ks = 0; ks < nks; ks += 2) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const double* wf = wl + (ks + h) * 4 * 9;
                const double* sr = sl + (ks + h) * 4 * NSR;
                double A[MT];
#pragma unroll
                for (int m = 0; m < MT; m++) A[m] = wf[tdx[m][0]] * wf[tdx[m][1]] * wf[tdx[m][2]];
#pragma unroll
                for (int n = 0; n < NCG; n++) {
                    const double B = sr[bi1[n]] * sr[bi2[n]];
                    if (n < NTF) {
#pragma unroll
                        for (int m = 0; m < MT; m++)
                            acc[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[m], B, acc[m][n], 0, 0, 0);
                    } else {
#pragma unroll
                        for (int m = 0; m < MT; m++)
                            accs[m][n - NTF] = __builtin_amdgcn_mfma_f64_4x4x4f64(A[m], B, accs[m][n - NTF], 0, 0, 0);
                    }
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }

    __device__ __forceinline__ void flush(double* grid, const Geom& g, int cell, int lane)
    {
        const int kx = __builtin_amdgcn_readfirstlane(cell % g.n[0]);
        const int ky = __builtin_amdgcn_readfirstlane((cell / g.n[0]) % g.n[1]);
        const int kz = __builtin_amdgcn_readfirstlane(cell / (g.n[0] * g.n[1]));
        auto wrap = [](int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); };
        auto target = [&](int t) {
            const int tx = wrap(kx + t % 3 - 1, g.n[0]);
            const int ty = wrap(ky + (t / 3) % 3 - 1, g.n[1]);
            const int tz = wrap(kz + t / 9 - 1, g.n[2]);
            return grid + (((int64_t)tz * g.n[1] + ty) * g.n[0] + tx) * QG;
        };
        // element r of lane l holds D[16m + 4r + l/16][16n + l%16]
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int t = 16 * m + 4 * r + (lane >> 4);
                if (t >= NT) continue;
                double* Gd = target(t);
#pragma unroll
                for (int n = 0; n < NTF; n++) unsafeAtomicAdd(&Gd[16 * n + (lane & 15)], acc[m][n][r]);
            }
        // 4x4x4 tiles: element of lane l is D[16m + 4((l >> 2) & 3) + l/16][16 NTF + 4k + l%4]
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int k = 0; k < NSM; k++) {
                const int t = 16 * m + 4 * ((lane >> 2) & 3) + (lane >> 4), q = 16 * NTF + 4 * k + (lane & 3);
                if (t >= NT || q >= QG) continue;
                unsafeAtomicAdd(&target(t)[q], accs[m][k]);
            }
    }
};

// Stage S of one step, in.Y = Y_S, grid = its deposit (rounded to T).
//  STORE 0 (the forward step): S = 1..3 write Y_{S+1} into ynext and the
//  accumulator into acc; S = 4 finishes acc = y_{n+1}, renormalizes it
//  (status deferred into err) and advances the particle's position.
//  STORE 1 (the VJP's recompute): S = 1..3 write Y_{S+1} into ynext only.
//  With DEP (S = 1..3), the stage also deposits Y_{S+1} onto gnext (its base
//  already there), the wave's particles on the matrix cores as the deposit
//  kernel does: positions are frozen within the step, so the cell's particles
//  are the ones whose stencils the deposit needs, and the stage state goes
//  from registers to the grid without a second pass over the particles.  That
//  pays at NF = 2; at NF = 3 the deposit's 36 columns and the gather's 36-value
//  rows in one kernel ran slower than the two kernels, which NF = 3 keeps.
template <int NF, int S, int STORE, typename T, bool DEP>
__global__ __launch_bounds__(64 * kCellWaves) void k_fwd_stage(AdjInT<T> in, Geom g, const int* start, const int* perm,
                                                               int64_t ncell, int bpx, const T* grid, HamParams hp,
                                                               const T* y0, T* ynext, T* acc, const T* Ll,
                                                               double maxError, double dt, double* pos_out,
                                                               unsigned long long* err, double* gnext)
{
    using WD = WaveDeposit<NF>;
    static_assert(!DEP || S < 4, "stage 4 deposits nothing");
    constexpr int nh = NH<NF>, S2 = 2 * nh;
    __shared__ double lds[DEP ? kCellWaves * WD::REG : 1];
    const int wave = threadIdx.x >> 6;
    double* Wf = lds + wave * WD::REG;
    double* Sr = Wf + 64 * 9;
    const int64_t np = in.np;
    const int cell = wave_cell(ncell, bpx);
    if (cell < 0) return;
    const int beg = __builtin_amdgcn_readfirstlane(start[cell]), end = __builtin_amdgcn_readfirstlane(start[cell + 1]);
    if (beg == end) return;
    const int lane = threadIdx.x & 63;
    const T a_next = T(S == 3 ? dt * 1.0 : dt * 0.5);
    const T a_acc = T(S == 1 || S == 4 ? dt * (1. / 6.) : dt * (1. / 3.));
    WD dep;
    if constexpr (DEP) dep.init(lane);
    for (int base = beg; base < end; base += 64) {
        const bool act = base + lane < end;
        const int64_t p = perm[act ? base + lane : beg];
        const double x[3] = {in.pos[p], in.pos[np + p], in.pos[2 * np + p]};
        double wd[3][3];
        T w[3][3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            int idx;
            shape3((x[d] - g.plo[d]) * g.dxi[d], g.order[d], idx, wd[d]);
#pragma unroll
            for (int e = 0; e < 3; e++) w[d][e] = T(wd[d][e]);
        }
        const T pt = in.pup[3 * np + p];
        const T v[3] = {in.pup[p] / pt, in.pup[np + p] / pt, in.pup[2 * np + p] / pt};
        T H[nh], Hb[nh];
        {
            const T sv[3] = {-v[0], -v[1], -v[2]};
            T inside[nh];
#pragma unroll
            for (int c = 0; c < nh; c++) inside[c] = T(0);
            gather_cell_box<NF, T>(g, cell, grid, w, sv, inside);
#pragma unroll
            for (int c = 0; c < nh; c++) {
                const T hv = T(hp.vac[c]) / pt, hs = T(hp.coef) * inside[c];
                H[c] = hv + hs;
                Hb[c] = is_im_comp<NF>(c) ? hs - hv : hv - hs;
            }
        }
        T q[nh]; // the sources of Y_{S+1} for the fused deposit (zero on inactive lanes)
#pragma unroll
        for (int c = 0; c < nh; c++) q[c] = T(0);
        if (act) {
            T f[nh], k[S2];
#pragma unroll
            for (int c = 0; c < nh; c++) f[c] = in.Y[(int64_t)c * np + p];
            commutator_rhs<NF, T>(H, f, k, T(1));
#pragma unroll
            for (int c = 0; c < nh; c++) f[c] = in.Y[(int64_t)(nh + c) * np + p];
            commutator_rhs<NF, T>(Hb, f, k + nh, T(1));
            if constexpr (S < 4) {
#pragma clang fp contract(off)
                T yv[S2];
#pragma unroll
                for (int c = 0; c < S2; c++) {
                    const T y = y0[(int64_t)c * np + p];
                    yv[c] = a_next * k[c] + y;
                    ynext[(int64_t)c * np + p] = yv[c];
                    if constexpr (STORE == 0)
                        acc[(int64_t)c * np + p] = a_acc * k[c] + (S == 1 ? y : acc[(int64_t)c * np + p]);
                }
                const T N = in.Nn[p], Nb = in.Nn[np + p];
#pragma unroll
                for (int c = 0; c < nh; c++)
                    q[c] = is_im_comp<NF>(c) ? N * yv[c] + Nb * yv[nh + c] : N * yv[c] - Nb * yv[nh + c];
            } else {
#pragma clang fp contract(off)
                // y_{n+1} in T, then Renormalize in FP64 (FlavoredNeutrinoContainer.cpp:89-113):
                // nu, then nubar unless nu failed; a failing half is left as it is
                double y[S2];
#pragma unroll
                for (int c = 0; c < S2; c++) y[c] = double(a_acc * k[c] + acc[(int64_t)c * np + p]);
                int e = renorm_fp64<NF>(y, double(Ll[p]), maxError);
                if (e == EMU_OK) e = renorm_fp64<NF>(y + nh, double(Ll[np + p]), maxError);
                if (e != EMU_OK) atomicMin(err, status_word(p, e));
#pragma unroll
                for (int c = 0; c < S2; c++) acc[(int64_t)c * np + p] = T(y[c]);
                double vv[3], nx[3];
#pragma unroll
                for (int d = 0; d < 3; d++)
                    vv[d] = double(in.pup[(int64_t)d * np + p]) / double(in.pup[3 * np + p]) * phys::c;
                int nc;
                end_position(x, vv, dt, g, nx, nc);
#pragma unroll
                for (int d = 0; d < 3; d++) pos_out[(int64_t)d * np + p] = nx[d];
            }
        }
        if constexpr (DEP) dep.add(Wf, Sr, lane, wd, q, v, end - base);
    }
    if constexpr (DEP) dep.flush(gnext, g, cell, lane);
}

// One reverse stage boundary in one pass, one cell per wave (s = 3 .. -1
// over the launches; G and C say which halves run):
//  G (stage s+1's gather back): lY_{s+1} = its commutator part (in lY) plus
//    the gather of gadj_{s+1} (N lq for nu, sigma Nbar lq for nubar), then
//    lam_out = lY_{s+1} + (lam at the first G, else lam_out) and
//    lam_Nbar += sigma fbar_{s+1} . lq;
//  C (stage s's commutator): lambda_k = alpha lam + beta lY_{s+1}, the
//    commutator adjoints (lY_s's commutator part, into lY in place) and mu_s,
//    which the wave deposits onto gadj_s on the f64 matrix cores (as
//    k_cell_deposit with KIND 1) without writing mu out.
// Against separate commutator, deposit and gather launches per stage, this
// form reads each particle once per boundary and holds one stage adjoint
// array instead of two and no mu array.  Every
// stencil width runs the 3 x 3 x 3 target set (the weights of an order-0
// dimension are (0, 1, 0), so its other targets receive zeros).
template <int NF, typename T, bool G, bool C>
__global__ __launch_bounds__(64 * kCellWaves) void k_vjp_fused(AdjInT<T> in, Geom g, const int* start,
                                                               const int* perm, int64_t ncell, int bpx,
                                                               const T* Ynext, const T* Ycur, const T* gadj_in,
                                                               const T* grid, double* gadj_out, HamParams hp,
                                                               const T* lam, T* lY, T* lam_out, double* lamNbar,
                                                               int first, double alpha, double beta)
{
    using WD = WaveDeposit<NF>;
    constexpr int nh = NH<NF>;
    __shared__ double lds[C ? kCellWaves * WD::REG : 1];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double* Wf = lds + wave * WD::REG;
    double* Sr = Wf + 64 * 9;
    const int cell = wave_cell(ncell, bpx);
    if (cell < 0) return;
    const int beg = __builtin_amdgcn_readfirstlane(start[cell]), end = __builtin_amdgcn_readfirstlane(start[cell + 1]);
    if (beg == end) return;
    const int64_t np = in.np;
    const T al = T(alpha), be = T(beta);
    WD dep;
    if constexpr (C) dep.init(lane);
    for (int base = beg; base < end; base += 64) {
        const bool act = base + lane < end;
        const int64_t p = perm[act ? base + lane : beg];
        // shape factors (FP64 for the deposit, T for the gathers) and direction
        const double x[3] = {in.pos[p], in.pos[np + p], in.pos[2 * np + p]};
        double wd[3][3];
        T w[3][3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            int idx;
            shape3((x[d] - g.plo[d]) * g.dxi[d], g.order[d], idx, wd[d]);
#pragma unroll
            for (int e = 0; e < 3; e++) w[d][e] = T(wd[d][e]);
        }
        const T pt = in.pup[3 * np + p];
        const T v[3] = {in.pup[p] / pt, in.pup[np + p] / pt, in.pup[2 * np + p] / pt};
        T lYn[2 * nh];
        if constexpr (G) {
            T lq[nh];
#pragma unroll
            for (int c = 0; c < nh; c++) lq[c] = T(0);
            gather_cell_box<NF, T>(g, cell, gadj_in, w, v, lq);
            if (act) {
                const T N = in.Nn[p], Nb = in.Nn[np + p];
                T ln = T(0);
#pragma unroll
                for (int c = 0; c < nh; c++) {
#pragma clang fp contract(off)
                    const T sg = is_im_comp<NF>(c) ? T(1) : T(-1);
                    const int64_t o0 = (int64_t)c * np + p, o1 = (int64_t)(nh + c) * np + p;
                    lYn[c] = lY[o0] + N * lq[c];
                    lYn[nh + c] = lY[o1] + sg * Nb * lq[c];
                    lam_out[o0] = T(1) * lYn[c] + (first ? lam[o0] : lam_out[o0]);
                    lam_out[o1] = T(1) * lYn[nh + c] + (first ? lam[o1] : lam_out[o1]);
                    ln += sg * Ynext[o1] * lq[c];
                }
                if (lamNbar) lamNbar[p] += double(ln);
            }
        }
        if constexpr (C) {
            const T sv[3] = {-v[0], -v[1], -v[2]};
            T H[nh], Hb[nh];
            {
                T inside[nh];
#pragma unroll
                for (int c = 0; c < nh; c++) inside[c] = T(0);
                gather_cell_box<NF, T>(g, cell, grid, w, sv, inside);
#pragma unroll
                for (int c = 0; c < nh; c++) {
                    const T hv = T(hp.vac[c]) / pt, hs = T(hp.coef) * inside[c];
                    H[c] = hv + hs;
                    Hb[c] = is_im_comp<NF>(c) ? hs - hv : hv - hs;
                }
            }
            T mu[nh];
#pragma unroll
            for (int c = 0; c < nh; c++) mu[c] = T(0);
            if (act) {
                T gH[2][nh];
#pragma unroll
                for (int bar = 0; bar < 2; bar++) {
                    T lk[nh], L[nh], f[nh], gy[nh];
#pragma unroll
                    for (int c = 0; c < nh; c++) {
#pragma clang fp contract(off)
                        const int64_t o = (int64_t)(bar * nh + c) * np + p;
                        const T t = al * lam[o];
                        if constexpr (G) lk[c] = be * lYn[bar * nh + c] + t;
                        else lk[c] = t;
                        f[c] = Ycur[o];
                    }
                    lam_matrix<NF, T>(lk, L);
                    commutator_rhs<NF, T>(L, bar ? Hb : H, gy, T(1)); // -i[Lam, A]
                    double_offdiag<NF, T>(gy);
#pragma unroll
                    for (int c = 0; c < nh; c++) lY[(int64_t)(bar * nh + c) * np + p] = gy[c];
                    commutator_rhs<NF, T>(f, L, gH[bar], T(1));       // -i[B, Lam]
                    double_offdiag<NF, T>(gH[bar]);
                }
#pragma unroll
                for (int c = 0; c < nh; c++)
                    mu[c] = T(hp.coef) * (gH[0][c] + (is_im_comp<NF>(c) ? gH[1][c] : -gH[1][c]));
            }
            // mu_s with its weights and (-v, 1) onto gadj_s
            dep.add(Wf, Sr, lane, wd, mu, sv, end - base);
        }
    }
    if constexpr (C) dep.flush(gadj_out, g, cell, lane);
}

__global__ void k_axpy(int64_t n, double a, const double* x, const double* y, double* out)
{
#pragma clang fp contract(off)
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        out[t] = a * x[t] + y[t];
}

__global__ void k_set_u64(unsigned long long* p, unsigned long long v)
{
    if (threadIdx.x == 0) *p = v;
}

__global__ void k_zero(double* p, int64_t n)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) p[t] = 0.0;
}

// wscale: the factor the particle weights N, Nbar carry (1, or a power of two
// in the FP32 form), which the potential's coefficient takes back out
HamParams ham_params(const emu_ctx* ctx, double wscale = 1.0)
{
    HamParams hp;
    const int nh = ctx->p.nf * ctx->p.nf;
    for (int c = 0; c < 9; c++) hp.vac[c] = c < nh ? ctx->M2[c] * phys::c4 / 2.0 / phys::hbar : 0.0;
    hp.coef = M_SQRT2 * phys::GF * ctx->g.inv_cell_volume / phys::hbar / wscale;
    return hp;
}

int64_t ncell_of(const emu_ctx* ctx) { return (int64_t)ctx->g.n[0] * ctx->g.n[1] * ctx->g.n[2]; }

} // namespace
} // namespace emu

using namespace emu;

// work: grid and grid adjoint [2][ncell][4 NH], mu [NH][np], the objective's
// [2 ncell + 8], then the cell sort: key, perm [np] ints, counts/start
// [ncell + 1], cursor [ncell] ints, and the scan's temporary storage.
static int64_t sort_offset_doubles(const emu_ctx* ctx, int64_t np)
{
    const int nh = ctx->p.nf * ctx->p.nf;
    return 2 * ncell_of(ctx) * 4 * nh + (int64_t)nh * np + 2 * ncell_of(ctx) + 8;
}

static size_t scan_bytes(const emu_ctx* ctx)
{
    size_t b = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, b, (const int*)nullptr, (int*)nullptr, (int)(ncell_of(ctx) + 1),
                                           ctx->stream);
    return b;
}

struct CellSort {
    int *key, *perm, *start, *cursor;
    void* tmp;
    size_t tmp_bytes;
};

static CellSort cell_sort_at(const emu_ctx* ctx, int64_t np, double* region)
{
    const int64_t nc = ncell_of(ctx);
    CellSort cs;
    int* base = (int*)region;
    cs.key = base;
    cs.perm = cs.key + np;
    cs.start = cs.perm + np;
    cs.cursor = cs.start + (nc + 1);
    const int64_t ints = 2 * np + 2 * nc + 1;
    const uintptr_t t = (uintptr_t)(region + (ints + 1) / 2);
    cs.tmp = (void*)((t + 255) & ~(uintptr_t)255);
    cs.tmp_bytes = scan_bytes(ctx);
    return cs;
}

// doubles of a cell sort region (ints, the scan's storage and its alignment)
static int64_t cell_sort_doubles(const emu_ctx* ctx, int64_t np)
{
    const int64_t ints = 2 * np + 2 * ncell_of(ctx) + 1;
    return (ints + 1) / 2 + 32 + (int64_t)(scan_bytes(ctx) + 7) / 8;
}

static CellSort cell_sort_space(const emu_ctx* ctx, int64_t np, double* work)
{
    const int64_t nc = ncell_of(ctx);
    CellSort cs;
    int* base = (int*)(work + sort_offset_doubles(ctx, np));
    cs.key = base;
    cs.perm = cs.key + np;
    cs.start = cs.perm + np;
    cs.cursor = cs.start + (nc + 1);
    const int64_t ints = 2 * np + 2 * nc + 1;
    // the scan's storage on a 256-byte boundary (the allocation reserves the slack)
    const uintptr_t t = (uintptr_t)(work + sort_offset_doubles(ctx, np) + (ints + 1) / 2);
    cs.tmp = (void*)((t + 255) & ~(uintptr_t)255);
    cs.tmp_bytes = scan_bytes(ctx);
    return cs;
}

extern "C" int64_t emu_adj_work_doubles(const emu_ctx* ctx, int64_t np)
{
    if (!ctx || np < 0) return -1;
    const int64_t nc = ncell_of(ctx);
    const int64_t ints = 2 * np + 2 * nc + 1;
    return sort_offset_doubles(ctx, np) + (ints + 1) / 2 + 32 + (int64_t)(scan_bytes(ctx) + 7) / 8;
}

#define ADJ_DISPATCH(kernel, ...)                                                                          \
    do {                                                                                                   \
        if (ctx->p.nf == 2) hipLaunchKernelGGL(kernel<2>, dim3(nblk(np)), dim3(kB), 0, ctx->stream, __VA_ARGS__); \
        else hipLaunchKernelGGL(kernel<3>, dim3(nblk(np)), dim3(kB), 0, ctx->stream, __VA_ARGS__);          \
    } while (0)

static double m00_of(const emu_ctx* ctx)
{
    return ctx->p.rho * ctx->p.Ye / phys::Mp / ctx->g.inv_cell_volume;
}

// blocks per XCD of a one-wave-per-cell launch
static int cell_blocks_per_xcd(int64_t ncell)
{
    const int64_t nb = (ncell + kCellWaves - 1) / kCellWaves;
    return (int)((nb + 7) / 8);
}

// k_cell_deposit for the stencil widths of the grid (3 where the shape order is
// above 0, else 1), one wave per cell of the cell list
template <int NF, int KIND, typename T>
static void cell_deposit_nf(emu_ctx* ctx, const AdjInT<T>& in, const CellSort& cs, const T* mu, double* grid)
{
    const int wx = ctx->g.order[0] > 0 ? 3 : 1, wy = ctx->g.order[1] > 0 ? 3 : 1, wz = ctx->g.order[2] > 0 ? 3 : 1;
    const int64_t nc = ncell_of(ctx);
    const int bpx = cell_blocks_per_xcd(nc);
    const int code = (wx == 3) * 4 + (wy == 3) * 2 + (wz == 3);
#define DEP_LAUNCH(X, Y, Z)                                                                                            \
    hipLaunchKernelGGL((k_cell_deposit<NF, X, Y, Z, KIND, T>), dim3(8 * bpx), dim3(64 * kCellWaves), 0, ctx->stream, in, \
                       ctx->g, cs.start, cs.perm, nc, bpx, mu, grid)
    switch (code) {
    case 0: DEP_LAUNCH(1, 1, 1); break;
    case 1: DEP_LAUNCH(1, 1, 3); break;
    case 2: DEP_LAUNCH(1, 3, 1); break;
    case 3: DEP_LAUNCH(1, 3, 3); break;
    case 4: DEP_LAUNCH(3, 1, 1); break;
    case 5: DEP_LAUNCH(3, 1, 3); break;
    case 6: DEP_LAUNCH(3, 3, 1); break;
    default: DEP_LAUNCH(3, 3, 3); break;
    }
#undef DEP_LAUNCH
}

template <int KIND, typename T>
static void cell_deposit(emu_ctx* ctx, const AdjInT<T>& in, const CellSort& cs, const T* mu, double* grid)
{
    if (in.np == 0) return;
    if (ctx->p.nf == 2) cell_deposit_nf<2, KIND, T>(ctx, in, cs, mu, grid);
    else cell_deposit_nf<3, KIND, T>(ctx, in, cs, mu, grid);
}

// the particles listed by the centre cell of their stencil (positions are
// frozen within a step, so every deposit of a call shares one list)
template <typename T> static int cell_sort(emu_ctx* ctx, const AdjInT<T>& in, const CellSort& cs)
{
    const int64_t np = in.np, nc = ncell_of(ctx);
    const int64_t zero_doubles = (2 * nc + 1 + 1) / 2; // start and cursor are adjacent ints
    hipLaunchKernelGGL(k_zero, dim3(nblk(zero_doubles)), dim3(kB), 0, ctx->stream, (double*)cs.start, zero_doubles);
    const unsigned nb = (unsigned)((np + kB - 1) / kB); // one thread per particle
    hipLaunchKernelGGL(k_cell_keys, dim3(nb), dim3(kB), 0, ctx->stream, np, in.pos, ctx->g, cs.key, cs.start);
    size_t bytes = cs.tmp_bytes;
    EMU_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(cs.tmp, bytes, cs.start, cs.start, (int)(nc + 1), ctx->stream));
    hipLaunchKernelGGL(k_cell_perm, dim3(nb), dim3(kB), 0, ctx->stream, np, cs.key, cs.start, cs.cursor, cs.perm);
    return EMU_OK;
}

// the deposit of Y onto the base grid (matter in component 0, carrying the
// weights' factor wscale like the deposited fields)
static void fill_base(emu_ctx* ctx, double* grid, double wscale)
{
    const int64_t ng = ncell_of(ctx) * 4 * ctx->p.nf * ctx->p.nf;
    const double m00 = m00_of(ctx) * wscale;
    if (ctx->p.nf == 2) hipLaunchKernelGGL(k_fill_base<2>, dim3(nblk(ng)), dim3(kB), 0, ctx->stream, grid, ncell_of(ctx), m00);
    else hipLaunchKernelGGL(k_fill_base<3>, dim3(nblk(ng)), dim3(kB), 0, ctx->stream, grid, ncell_of(ctx), m00);
}

template <typename T>
static void deposit(emu_ctx* ctx, const AdjInT<T>& in, const CellSort& cs, double* grid, double wscale = 1.0)
{
    fill_base(ctx, grid, wscale);
    cell_deposit<0, T>(ctx, in, cs, (const T*)nullptr, grid);
}

extern "C" int emu_adj_rhs(emu_ctx* ctx, int64_t np, const double* pos, const double* pup, const double* Nn,
                           const double* Y, double* k, double* work)
{
    if (!ctx || np < 0 || !work) return EMU_ERR_ARG;
    if (np == 0) return EMU_OK;
    const AdjIn in{np, pos, pup, Nn, Y};
    double* grid = work;
    const CellSort cs = cell_sort_space(ctx, np, work);
    if (int e = cell_sort(ctx, in, cs)) return e;
    deposit(ctx, in, cs, grid);
    ADJ_DISPATCH(k_rhs, in, ctx->g, cs.perm, grid, ham_params(ctx), k);
    EMU_LAUNCH_CHECK(ctx, "adj_rhs");
    return EMU_OK;
}

extern "C" int emu_adj_rhs_vjp(emu_ctx* ctx, int64_t np, const double* pos, const double* pup, const double* Nn,
                               const double* Y, const double* lam_k, double* lam_Y, double* lam_Nbar, double* work)
{
    if (!ctx || np < 0 || !work) return EMU_ERR_ARG;
    if (np == 0) return EMU_OK;
    const AdjIn in{np, pos, pup, Nn, Y};
    const int nh = ctx->p.nf * ctx->p.nf;
    const int64_t ng = ncell_of(ctx) * 4 * nh;
    double* grid = work;
    double* gadj = work + ng;
    double* mu = gadj + ng;
    const CellSort cs = cell_sort_space(ctx, np, work);
    if (int e = cell_sort(ctx, in, cs)) return e;
    deposit(ctx, in, cs, grid);
    ADJ_DISPATCH(k_vjp_comm, in, ctx->g, cs.perm, grid, ham_params(ctx), lam_k, lam_Y, mu);
    hipLaunchKernelGGL(k_zero, dim3(nblk(ng)), dim3(kB), 0, ctx->stream, gadj, ng);
    cell_deposit<1, double>(ctx, in, cs, mu, gadj);
    ADJ_DISPATCH(k_vjp_gather, in, ctx->g, cs.perm, gadj, lam_Y, lam_Nbar);
    EMU_LAUNCH_CHECK(ctx, "adj_rhs_vjp");
    return EMU_OK;
}

template <typename T>
static int adj_objective(emu_ctx* ctx, int64_t np, const double* pos, const T* Nn, const T* Y, double* J, T* dJdY,
                         double* work)
{
    if (!ctx || np < 0 || !work || !J) return EMU_ERR_ARG;
    const AdjInT<T> in{np, pos, nullptr, Nn, Y};
    const int64_t nc = ncell_of(ctx);
    double* g2 = work; // [ncell][2] N_01, then J
    double* dJ = g2 + 2 * nc;
    hipLaunchKernelGGL(k_zero, dim3(nblk(2 * nc + 1)), dim3(kB), 0, ctx->stream, g2, 2 * nc + 1);
    if (np > 0) {
        if (ctx->p.nf == 2) hipLaunchKernelGGL((k_obj_deposit<2, T>), dim3(nblk(np)), dim3(kB), 0, ctx->stream, in, ctx->g, g2);
        else hipLaunchKernelGGL((k_obj_deposit<3, T>), dim3(nblk(np)), dim3(kB), 0, ctx->stream, in, ctx->g, g2);
    }
    hipLaunchKernelGGL(k_obj_sum, dim3(nblk(nc)), dim3(kB), 0, ctx->stream, g2, nc, dJ);
    if (dJdY && np > 0) {
        if (ctx->p.nf == 2)
            hipLaunchKernelGGL((k_obj_grad<2, T>), dim3(nblk(np)), dim3(kB), 0, ctx->stream, in, ctx->g, g2, dJdY);
        else hipLaunchKernelGGL((k_obj_grad<3, T>), dim3(nblk(np)), dim3(kB), 0, ctx->stream, in, ctx->g, g2, dJdY);
    }
    EMU_LAUNCH_CHECK(ctx, "adj_objective");
    EMU_HIP(ctx, hipMemcpyAsync(J, dJ, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    EMU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return EMU_OK;
}

extern "C" int emu_adj_objective(emu_ctx* ctx, int64_t np, const double* pos, const double* Nn, const double* Y,
                                 double* J, double* dJdY, double* work)
{
    return adj_objective<double>(ctx, np, pos, Nn, Y, J, dJdY, work);
}

extern "C" int emu_adj_objective_f32(emu_ctx* ctx, int64_t np, const double* pos, const float* Nn, const float* Y,
                                     double* J, float* dJdY, double* work)
{
    return adj_objective<float>(ctx, np, pos, Nn, Y, J, dJdY, work);
}

extern "C" int emu_adj_advance_positions(emu_ctx* ctx, int64_t np, double* pos, const double* pup, double dt)
{
    if (!ctx || np < 0) return EMU_ERR_ARG;
    if (np == 0) return EMU_OK;
    hipLaunchKernelGGL(k_positions, dim3(nblk(np)), dim3(kB), 0, ctx->stream, np, pos, pup, dt, ctx->g);
    EMU_LAUNCH_CHECK(ctx, "adj_advance_positions");
    return EMU_OK;
}

extern "C" int emu_adj_renormalize(emu_ctx* ctx, int64_t np, double* Y, const double* Ll, int64_t* bad_index)
{
    if (!ctx || np < 0) return EMU_ERR_ARG;
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, ctx->stream, ctx->d_red + 6, ~0ull);
    if (np > 0) ADJ_DISPATCH(k_renorm_soa, np, Y, Ll, ctx->p.maxError, ctx->d_red + 6);
    EMU_LAUNCH_CHECK(ctx, "adj_renormalize");
    unsigned long long h;
    EMU_HIP(ctx, hipMemcpyAsync(&h, ctx->d_red + 6, 8, hipMemcpyDeviceToHost, ctx->stream));
    EMU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return status_decode(h, bad_index);
}

extern "C" int emu_adj_axpy(emu_ctx* ctx, int64_t n, double a, const double* x, const double* y, double* out)
{
    if (!ctx || n < 0) return EMU_ERR_ARG;
    if (n == 0) return EMU_OK;
    hipLaunchKernelGGL(k_axpy, dim3(nblk(n)), dim3(kB), 0, ctx->stream, n, a, x, y, out);
    EMU_LAUNCH_CHECK(ctx, "adj_axpy");
    return EMU_OK;
}

// ---------------------------------------------------------------------------
// Step-level calls, in FP64 (T = double) or in the FP32-bulk form (T = float:
// the flavour arrays, the particle constants and the adjoints in FP32, with
// the gathers, commutators and stage algebra in FP32; the deposits sum in
// FP64 and Renormalize runs in FP64, as in the fused FP32-bulk step).
// Work layout, in doubles: the four stage grids [4][ncell][4 NH] and two
// grid adjoints (FP64); for T = float two T copies of grids, which the
// gathers read; then, in T, the stage states [3][2 NH][np] (Y2, Y3, Y4) and
// one stage adjoint [2 NH][np]; then the cell sort.
// ---------------------------------------------------------------------------
namespace {
template <typename T> int64_t t_doubles(int64_t n) { return (n * (int64_t)sizeof(T) + 7) / 8; }

template <typename T> struct StepWork {
    double* G[4];
    double* gadj[2];
    T* gT[2]; // the grids a float gather reads (the double gathers read the grids themselves)
    T* Ys[3];
    T* lY;
    CellSort cs;
};

template <typename T> StepWork<T> step_work(const emu_ctx* ctx, int64_t np, double* work)
{
    const int nh = ctx->p.nf * ctx->p.nf;
    const int64_t ng = ncell_of(ctx) * 4 * nh, ns = 2 * nh * np;
    StepWork<T> w;
    double* b = work;
    for (int s = 0; s < 4; s++, b += ng) w.G[s] = b;
    for (int s = 0; s < 2; s++, b += ng) w.gadj[s] = b;
    w.gT[0] = w.gT[1] = nullptr;
    if constexpr (sizeof(T) != sizeof(double)) {
        for (int s = 0; s < 2; s++, b += t_doubles<T>(ng)) w.gT[s] = (T*)b;
    }
    for (int s = 0; s < 3; s++, b += t_doubles<T>(ns)) w.Ys[s] = (T*)b;
    w.lY = (T*)b;
    b += t_doubles<T>(ns);
    w.cs = cell_sort_at(ctx, np, b);
    return w;
}

template <typename T> int64_t step_work_doubles(const emu_ctx* ctx, int64_t np)
{
    const int nh = ctx->p.nf * ctx->p.nf;
    const int64_t ng = ncell_of(ctx) * 4 * nh, ns = 2 * nh * np;
    return 6 * ng + (sizeof(T) != sizeof(double) ? 2 * t_doubles<T>(ng) : 0) + 4 * t_doubles<T>(ns) +
           cell_sort_doubles(ctx, np);
}

__global__ void k_to_float(const double* in, float* out, int64_t n)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        out[t] = (float)in[t];
}

// The grid a gather reads: the FP64 grid, or its FP32 rounding in buffer
// `slot` (the FP32-bulk gather's input, as the fused step rounds its grid
// when staging it).
template <typename T> const T* gather_grid(emu_ctx* ctx, const StepWork<T>& w, const double* g, int slot = 0)
{
    if constexpr (sizeof(T) == sizeof(double)) {
        return g;
    } else {
        const int64_t ng = ncell_of(ctx) * 4 * ctx->p.nf * ctx->p.nf;
        hipLaunchKernelGGL(k_to_float, dim3(nblk(ng)), dim3(kB), 0, ctx->stream, g, w.gT[slot], ng);
        return w.gT[slot];
    }
}

int arm_status(emu_ctx* ctx)
{
    if (!ctx->adj_status_armed) {
        hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, ctx->stream, ctx->d_red + 5, ~0ull);
        ctx->adj_status_armed = true;
    }
    return EMU_OK;
}

// Stage S reads the deposit of Y_S (grid) and, for S < 4, leaves the deposit
// of Y_{S+1} on gnext: within the stage kernel at NF = 2, as a kernel of its
// own after it at NF = 3.
template <int NF, int S, int STORE, typename T>
void fwd_stage(emu_ctx* ctx, const AdjInT<T>& in, const StepWork<T>& w, const double* grid, const T* y0, T* ynext,
               T* acc, const T* Ll, double dt, double* pos, double wscale, double* gnext)
{
    constexpr bool DEP = S < 4 && NF == 2;
    const int64_t nc = ncell_of(ctx);
    const int bpx = cell_blocks_per_xcd(nc);
    if (DEP) fill_base(ctx, gnext, wscale);
    const T* gg = gather_grid<T>(ctx, w, grid);
    hipLaunchKernelGGL((k_fwd_stage<NF, S, STORE, T, DEP>), dim3(8 * bpx), dim3(64 * kCellWaves), 0, ctx->stream, in,
                       ctx->g, w.cs.start, w.cs.perm, nc, bpx, gg, ham_params(ctx, wscale), y0, ynext, acc, Ll,
                       ctx->p.maxError, dt, pos, ctx->d_red + 5, gnext);
    if (S < 4 && !DEP) {
        AdjInT<T> next = in;
        next.Y = ynext;
        deposit(ctx, next, w.cs, gnext, wscale);
    }
}

// A step's kept data (emu_adj_step_keep): the stage states Y2, Y3, Y4
// ([3][2 NH][np], in T) and the four stage deposits ([4][ng], FP64).
template <typename T> struct StepKeep {
    T* Ys[3];
    double* G[4];
};

template <typename T> StepKeep<T> step_keep(const emu_ctx* ctx, int64_t np, double* keep)
{
    const int nh = ctx->p.nf * ctx->p.nf;
    const int64_t ng = ncell_of(ctx) * 4 * nh, ns = 2 * nh * np;
    StepKeep<T> k;
    double* b = keep;
    for (int s = 0; s < 3; s++, b += t_doubles<T>(ns)) k.Ys[s] = (T*)b;
    for (int s = 0; s < 4; s++, b += ng) k.G[s] = b;
    return k;
}

template <typename T> int64_t step_keep_doubles(const emu_ctx* ctx, int64_t np)
{
    const int nh = ctx->p.nf * ctx->p.nf;
    return 3 * t_doubles<T>(2 * nh * np) + 4 * ncell_of(ctx) * 4 * nh;
}

// Stage s reads Y_s and its deposit G[s - 1] and writes Y_{s+1} into Ys[s - 1]:
// every stage state and deposit has its own array, so that a step run with
// `keep` leaves them all for its VJP.
template <int NF, typename T>
int adj_step_nf(emu_ctx* ctx, int64_t np, double* pos, const T* pup, const T* Nn, const T* Ll, const T* y_in,
                T* y_out, double dt, double wscale, double* work, double* keep)
{
    StepWork<T> w = step_work<T>(ctx, np, work);
    if (keep) {
        const StepKeep<T> k = step_keep<T>(ctx, np, keep);
        for (int s = 0; s < 3; s++) w.Ys[s] = k.Ys[s];
        for (int s = 0; s < 4; s++) w.G[s] = k.G[s];
    }
    AdjInT<T> in{np, pos, pup, Nn, y_in};
    if (int e = cell_sort(ctx, in, w.cs)) return e;
    deposit(ctx, in, w.cs, w.G[0], wscale);
    fwd_stage<NF, 1, 0, T>(ctx, in, w, w.G[0], y_in, w.Ys[0], y_out, Ll, dt, pos, wscale, w.G[1]);
    in.Y = w.Ys[0];
    fwd_stage<NF, 2, 0, T>(ctx, in, w, w.G[1], y_in, w.Ys[1], y_out, Ll, dt, pos, wscale, w.G[2]);
    in.Y = w.Ys[1];
    fwd_stage<NF, 3, 0, T>(ctx, in, w, w.G[2], y_in, w.Ys[2], y_out, Ll, dt, pos, wscale, w.G[3]);
    in.Y = w.Ys[2];
    fwd_stage<NF, 4, 0, T>(ctx, in, w, w.G[3], y_in, (T*)nullptr, y_out, Ll, dt, pos, wscale, nullptr);
    return EMU_OK;
}

template <int NF, typename T>
int adj_step_vjp_nf(emu_ctx* ctx, int64_t np, const double* pos, const T* pup, const T* Nn, const T* y_in,
                    const T* lam_in, T* lam_out, double* lam_Nbar, double dt, double wscale, double* work,
                    const double* keep)
{
    const int nh = NF * NF;
    const int64_t ng = ncell_of(ctx) * 4 * nh;
    StepWork<T> w = step_work<T>(ctx, np, work);
    AdjInT<T> in{np, pos, pup, Nn, y_in};
    if (int e = cell_sort(ctx, in, w.cs)) return e;
    if (keep) {
        // the forward step's own stage states and deposits
        const StepKeep<T> k = step_keep<T>(ctx, np, (double*)keep);
        for (int s = 0; s < 3; s++) w.Ys[s] = k.Ys[s];
        for (int s = 0; s < 4; s++) w.G[s] = k.G[s];
    } else {
        // recompute the stage states and their deposits
        deposit(ctx, in, w.cs, w.G[0], wscale);
        fwd_stage<NF, 1, 1, T>(ctx, in, w, w.G[0], y_in, w.Ys[0], nullptr, nullptr, dt, nullptr, wscale, w.G[1]);
        in.Y = w.Ys[0];
        fwd_stage<NF, 2, 1, T>(ctx, in, w, w.G[1], y_in, w.Ys[1], nullptr, nullptr, dt, nullptr, wscale, w.G[2]);
        in.Y = w.Ys[1];
        fwd_stage<NF, 3, 1, T>(ctx, in, w, w.G[2], y_in, w.Ys[2], nullptr, nullptr, dt, nullptr, wscale, w.G[3]);
    }
    const T* Yst[4] = {y_in, w.Ys[0], w.Ys[1], w.Ys[2]};
    // reverse: lambda_k4 = dt/6 lam; lambda_k3 = dt/3 lam + dt lY4;
    // lambda_k2 = dt/3 lam + dt/2 lY3; lambda_k1 = dt/6 lam + dt/2 lY2;
    // lam_out = lam + lY4 + lY3 + lY2 + lY1.  Launch s (3 .. -1) finishes
    // stage s+1's lY (the gather back of its grid adjoint) and runs stage s's
    // commutator adjoints, depositing mu_s onto the other grid adjoint.
    const int bpx = cell_blocks_per_xcd(ncell_of(ctx));
    const double alpha[4] = {dt * (1. / 6.), dt * (1. / 3.), dt * (1. / 3.), dt * (1. / 6.)};
    const double beta[4] = {dt * 0.5, dt * 0.5, dt * 1.0, 0.0};
    const HamParams hp = ham_params(ctx, wscale);
    const dim3 grid_dim(8 * bpx), block_dim(64 * kCellWaves);
    for (int s = 3; s >= -1; s--) {
        const bool g_half = s < 3, c_half = s >= 0;
        double* gout = c_half ? w.gadj[s & 1] : nullptr;
        if (c_half) hipLaunchKernelGGL(k_zero, dim3(nblk(ng)), dim3(kB), 0, ctx->stream, gout, ng);
        const T* gin = g_half ? gather_grid<T>(ctx, w, w.gadj[(s + 1) & 1], 1) : nullptr;
        const T* gs = c_half ? gather_grid<T>(ctx, w, w.G[s], 0) : nullptr;
        const T* ynext = g_half ? Yst[s + 1] : nullptr;
        const T* ycur = c_half ? Yst[s] : nullptr;
        const double a = c_half ? alpha[s] : 0.0, b = c_half ? beta[s] : 0.0;
        const int first = s == 2 ? 1 : 0;
        if (!g_half)
            hipLaunchKernelGGL((k_vjp_fused<NF, T, false, true>), grid_dim, block_dim, 0, ctx->stream, in, ctx->g,
                               w.cs.start, w.cs.perm, ncell_of(ctx), bpx, ynext, ycur, gin, gs, gout, hp, lam_in, w.lY,
                               lam_out, lam_Nbar, first, a, b);
        else if (c_half)
            hipLaunchKernelGGL((k_vjp_fused<NF, T, true, true>), grid_dim, block_dim, 0, ctx->stream, in, ctx->g,
                               w.cs.start, w.cs.perm, ncell_of(ctx), bpx, ynext, ycur, gin, gs, gout, hp, lam_in, w.lY,
                               lam_out, lam_Nbar, first, a, b);
        else
            hipLaunchKernelGGL((k_vjp_fused<NF, T, true, false>), grid_dim, block_dim, 0, ctx->stream, in, ctx->g,
                               w.cs.start, w.cs.perm, ncell_of(ctx), bpx, ynext, ycur, gin, gs, gout, hp, lam_in, w.lY,
                               lam_out, lam_Nbar, first, a, b);
    }
    return EMU_OK;
}

template <typename T>
int adj_step_any(emu_ctx* ctx, int64_t np, double* pos, const T* pup, const T* Nn, const T* Ll, const T* y_in,
                 T* y_out, double dt, double wscale, double* work, double* keep)
{
    if (!ctx || np < 0 || !work || y_in == y_out || !(wscale > 0.0)) return EMU_ERR_ARG;
    arm_status(ctx);
    if (np == 0) return EMU_OK;
    const int e = ctx->p.nf == 2 ? adj_step_nf<2, T>(ctx, np, pos, pup, Nn, Ll, y_in, y_out, dt, wscale, work, keep)
                                 : adj_step_nf<3, T>(ctx, np, pos, pup, Nn, Ll, y_in, y_out, dt, wscale, work, keep);
    if (e) return e;
    EMU_LAUNCH_CHECK(ctx, "adj_step");
    return EMU_OK;
}

template <typename T>
int adj_step_vjp_any(emu_ctx* ctx, int64_t np, const double* pos, const T* pup, const T* Nn, const T* y_in,
                     const T* lam_in, T* lam_out, double* lam_Nbar, double dt, double wscale, double* work,
                     const double* keep)
{
    if (!ctx || np < 0 || !work || lam_in == lam_out || !(wscale > 0.0)) return EMU_ERR_ARG;
    if (np == 0) return EMU_OK;
    const int e = ctx->p.nf == 2 ? adj_step_vjp_nf<2, T>(ctx, np, pos, pup, Nn, y_in, lam_in, lam_out, lam_Nbar, dt,
                                                         wscale, work, keep)
                                 : adj_step_vjp_nf<3, T>(ctx, np, pos, pup, Nn, y_in, lam_in, lam_out, lam_Nbar, dt,
                                                         wscale, work, keep);
    if (e) return e;
    EMU_LAUNCH_CHECK(ctx, "adj_step_vjp");
    return EMU_OK;
}
} // namespace

extern "C" int64_t emu_adj_step_work_doubles(const emu_ctx* ctx, int64_t np)
{
    if (!ctx || np < 0) return -1;
    return step_work_doubles<double>(ctx, np);
}

extern "C" int emu_adj_step(emu_ctx* ctx, int64_t np, double* pos, const double* pup, const double* Nn,
                            const double* Ll, const double* y_in, double* y_out, double dt, double* work)
{
    return adj_step_any<double>(ctx, np, pos, pup, Nn, Ll, y_in, y_out, dt, 1.0, work, nullptr);
}

extern "C" int emu_adj_step_vjp(emu_ctx* ctx, int64_t np, const double* pos, const double* pup, const double* Nn,
                                const double* y_in, const double* lam_in, double* lam_out, double* lam_Nbar, double dt,
                                double* work)
{
    return adj_step_vjp_any<double>(ctx, np, pos, pup, Nn, y_in, lam_in, lam_out, lam_Nbar, dt, 1.0, work, nullptr);
}

extern "C" int64_t emu_adj_step_keep_doubles(const emu_ctx* ctx, int64_t np)
{
    if (!ctx || np < 0) return -1;
    return step_keep_doubles<double>(ctx, np);
}

extern "C" int emu_adj_step_keep(emu_ctx* ctx, int64_t np, double* pos, const double* pup, const double* Nn,
                                 const double* Ll, const double* y_in, double* y_out, double dt, double* work,
                                 double* keep)
{
    if (!keep) return EMU_ERR_ARG;
    return adj_step_any<double>(ctx, np, pos, pup, Nn, Ll, y_in, y_out, dt, 1.0, work, keep);
}

extern "C" int emu_adj_step_vjp_kept(emu_ctx* ctx, int64_t np, const double* pos, const double* pup,
                                     const double* Nn, const double* y_in, const double* keep, const double* lam_in,
                                     double* lam_out, double* lam_Nbar, double dt, double* work)
{
    if (!keep) return EMU_ERR_ARG;
    return adj_step_vjp_any<double>(ctx, np, pos, pup, Nn, y_in, lam_in, lam_out, lam_Nbar, dt, 1.0, work, keep);
}

extern "C" int64_t emu_adj_step_work_doubles_f32(const emu_ctx* ctx, int64_t np)
{
    if (!ctx || np < 0) return -1;
    return step_work_doubles<float>(ctx, np);
}

extern "C" int emu_adj_step_f32(emu_ctx* ctx, int64_t np, double* pos, const float* pup, const float* Nn,
                                const float* Ll, const float* y_in, float* y_out, double dt, double wscale,
                                double* work)
{
    return adj_step_any<float>(ctx, np, pos, pup, Nn, Ll, y_in, y_out, dt, wscale, work, nullptr);
}

extern "C" int emu_adj_step_vjp_f32(emu_ctx* ctx, int64_t np, const double* pos, const float* pup, const float* Nn,
                                    const float* y_in, const float* lam_in, float* lam_out, double* lam_Nbar,
                                    double dt, double wscale, double* work)
{
    return adj_step_vjp_any<float>(ctx, np, pos, pup, Nn, y_in, lam_in, lam_out, lam_Nbar, dt, wscale, work,
                                   nullptr);
}

extern "C" int64_t emu_adj_step_keep_doubles_f32(const emu_ctx* ctx, int64_t np)
{
    if (!ctx || np < 0) return -1;
    return step_keep_doubles<float>(ctx, np);
}

extern "C" int emu_adj_step_keep_f32(emu_ctx* ctx, int64_t np, double* pos, const float* pup, const float* Nn,
                                     const float* Ll, const float* y_in, float* y_out, double dt, double wscale,
                                     double* work, double* keep)
{
    if (!keep) return EMU_ERR_ARG;
    return adj_step_any<float>(ctx, np, pos, pup, Nn, Ll, y_in, y_out, dt, wscale, work, keep);
}

extern "C" int emu_adj_step_vjp_kept_f32(emu_ctx* ctx, int64_t np, const double* pos, const float* pup,
                                         const float* Nn, const float* y_in, const double* keep, const float* lam_in,
                                         float* lam_out, double* lam_Nbar, double dt, double wscale, double* work)
{
    if (!keep) return EMU_ERR_ARG;
    return adj_step_vjp_any<float>(ctx, np, pos, pup, Nn, y_in, lam_in, lam_out, lam_Nbar, dt, wscale, work, keep);
}

namespace {
// out[r][i] = in[r][idx[i]] (gather) or out[r][idx[i]] = in[r][i] (scatter;
// accumulate adds), rows of np doubles
template <typename T>
__global__ void k_permute_rows(int rows, int64_t np, const T* in, const int* idx, T* out, int mode)
{
    const int64_t n = (int64_t)rows * np;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = t / np, i = t - r * np;
        const int64_t j = idx[i];
        if (mode == 0) out[t] = in[r * np + j];
        else if (mode == 1) out[r * np + j] = in[t];
        else out[r * np + j] += in[t];
    }
}

__global__ void k_copy_ints(int64_t n, const int* in, int* out)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        out[t] = in[t];
}
} // namespace

extern "C" int emu_adj_cell_order(emu_ctx* ctx, int64_t np, const double* pos, int* order, double* work)
{
    if (!ctx || np < 0 || !work || (np > 0 && (!pos || !order))) return EMU_ERR_ARG;
    if (np == 0) return EMU_OK;
    const CellSort cs = cell_sort_at(ctx, np, work);
    const AdjIn in{np, pos, nullptr, nullptr, nullptr};
    if (int e = cell_sort(ctx, in, cs)) return e;
    hipLaunchKernelGGL(k_copy_ints, dim3(nblk(np)), dim3(kB), 0, ctx->stream, np, cs.perm, order);
    EMU_LAUNCH_CHECK(ctx, "adj_cell_order");
    return EMU_OK;
}

template <typename T>
static int adj_permute(emu_ctx* ctx, int rows, int64_t np, const T* in, const int* idx, T* out, int mode)
{
    if (!ctx || rows < 0 || np < 0 || mode < 0 || mode > 2 || in == out) return EMU_ERR_ARG;
    if (rows == 0 || np == 0) return EMU_OK;
    hipLaunchKernelGGL(k_permute_rows<T>, dim3(nblk((int64_t)rows * np)), dim3(kB), 0, ctx->stream, rows, np, in, idx,
                       out, mode);
    EMU_LAUNCH_CHECK(ctx, "adj_permute");
    return EMU_OK;
}

extern "C" int emu_adj_permute(emu_ctx* ctx, int rows, int64_t np, const double* in, const int* idx, double* out,
                               int mode)
{
    return adj_permute<double>(ctx, rows, np, in, idx, out, mode);
}

extern "C" int emu_adj_permute_f32(emu_ctx* ctx, int rows, int64_t np, const float* in, const int* idx, float* out,
                                   int mode)
{
    return adj_permute<float>(ctx, rows, np, in, idx, out, mode);
}

extern "C" int emu_adj_status(emu_ctx* ctx, int64_t* bad_index)
{
    if (!ctx) return EMU_ERR_ARG;
    arm_status(ctx);
    unsigned long long h;
    EMU_HIP(ctx, hipMemcpyAsync(&h, ctx->d_red + 5, 8, hipMemcpyDeviceToHost, ctx->stream));
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, ctx->stream, ctx->d_red + 5, ~0ull);
    EMU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return status_decode(h, bad_index);
}
