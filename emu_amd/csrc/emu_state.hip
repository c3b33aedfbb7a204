// emu_state.hip -- the fused RK4 step on a library-owned, cell-sorted SoA
// particle state (the integrator loop of main.cpp:129-213).
//
// Per step (stage s reads the grid deposited from the stage-s state):
//   stage 1..3 kernels (emu_stage.h): gather + Hamiltonian + commutator + RK
//     update of the particle, then deposit of the next stage state into the
//     other grid; stage 1 also bins each particle's end-of-step position;
//   a stable counting sort plan by the new cell (ranks and bases);
//   compute_dt reduction over the stage-4 deposit (main.cpp:193-197);
//   stage 4 kernel: y_{n+1}, positions, periodic shift, Renormalize (FP64),
//     each particle written straight to its slot in the new cell order;
//   the deposit of y_{n+1} that the next step's stage 1 reads.
//
// The mesh holds, per cell and flavour component c, the combined fields
//   M_c  = N_c - conj(Nbar)_c  (+ rho Ye/Mp dV on 00, the matter term)
//   Phi_d,c = F_d,c - conj(Fbar_d,c),  d = x, y, z
// which is all the Hamiltonian needs (H_SI = sqrt2 GF/dV (M - p.Phi),
// Hbar = conj(H_vac - H_SI - H_matter), generate_code.py:296-357); compute_dt
// additionally needs N_ii, Nbar_ii and Im(Fbar_d) off the diagonal, which the
// stage-3 kernel deposits into an auxiliary grid.
#include "emu_internal.h"
#include "emu_stage.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace emu {
namespace {

// Base value of the deposit grid: zero except the matter term rho Ye/Mp dV
// on M_00 (inside_parentheses of generate_code.py:335-342, moved into the
// field because it enters the gather with the same weights).
__global__ void k_fill_grid(double* g, int64_t ncell, int Q, double m00, double* aux, int naux)
{
    const int64_t n = ncell * Q;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        g[t] = (t % Q) == 0 ? m00 : 0.0;
    if (aux) {
        const int64_t na = ncell * naux;
        for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < na; t += (int64_t)gridDim.x * blockDim.x)
            aux[t] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// Stable counting sort by cell after each step.  A particle moves at most one
// cell per dimension per step (c dt < dx, checked on the host), so its move
// is one of 27 offsets e.  Particles arriving in cell t are ordered by the
// canonical index of e, then by their position in their old cell: every
// particle's destination is base[old cell][e] + (its rank among the particles
// of its old cell with the same e), which makes the order deterministic.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int move1(int from, int to, int n, int* bad)
{
    const int dd = pmod(to - from, n);
    if (dd == 0) return 0;
    if (dd == 1) return 1;
    if (dd == n - 1) return -1;
    *bad = 1;
    return 0;
}

// One wave per old cell: move index and rank within (cell, move); m[c][e].
// A move of more than one cell (only a particle with |p| > p_t can outrun
// c dt < dx) is reported as EMU_ERR_MOVE with the particle's id.
__global__ void k_move_ranks(const int* __restrict__ newcell, const int* __restrict__ cell_start,
                             const int* __restrict__ cell_cnt, const int64_t* __restrict__ ids, Geom g,
                             int* __restrict__ rank, unsigned char* __restrict__ mv, int* __restrict__ m,
                             unsigned long long* err)
{
    const int c = blockIdx.x;
    const int lane = threadIdx.x;
    const int cx = c % g.n[0], cy = (c / g.n[0]) % g.n[1], cz = c / (g.n[0] * g.n[1]);
    const int beg = cell_start[c], end = beg + cell_cnt[c];
    int running = 0; // lane e (< 27) holds the running count of move e
    for (int base = beg; base < end; base += 64) {
        const int ip = base + lane;
        const bool act = ip < end;
        int e = 27;
        if (act) {
            const int t = newcell[ip];
            const int tx = t % g.n[0], ty = (t / g.n[0]) % g.n[1], tz = t / (g.n[0] * g.n[1]);
            int bad = 0;
            const int ex = move1(cx, tx, g.n[0], &bad), ey = move1(cy, ty, g.n[1], &bad),
                      ez = move1(cz, tz, g.n[2], &bad);
            if (bad) atomicMin(err, status_word(ids[ip], EMU_ERR_MOVE));
            e = (ez + 1) * 9 + (ey + 1) * 3 + (ex + 1);
        }
        // one pass per move present in the chunk (most particles stay in their
        // cell, so there are few): the lanes of move q get ranks after the
        // earlier chunks' particles of that move, in lane order
        int myrank = 0;
        unsigned long long left = __ballot(act);
        while (left) {
            const int q = __shfl(e, __ffsll((long long)left) - 1);
            const unsigned long long msk = __ballot(act && e == q);
            const int before = __shfl(running, q);
            if (act && e == q) myrank = before + __popcll(msk & ((1ull << lane) - 1ull));
            if (lane == q) running += __popcll(msk);
            left &= ~msk;
        }
        if (act) {
            rank[ip] = myrank;
            mv[ip] = (unsigned char)e;
        }
    }
    if (lane < 27) m[(int64_t)c * 27 + lane] = running;
}

__device__ __forceinline__ bool valid_move(int e, const Geom& g)
{
    const int ex = e % 3 - 1, ey = (e / 3) % 3 - 1, ez = e / 9 - 1;
    const int ed[3] = {ex, ey, ez};
    for (int d = 0; d < 3; d++) {
        if (g.n[d] == 1 && ed[d] != 0) return false;
        if (g.n[d] == 2 && ed[d] == -1) return false;
    }
    return true;
}

__device__ __forceinline__ int64_t source_of(int t, int e, const Geom& g)
{
    const int tx = t % g.n[0], ty = (t / g.n[0]) % g.n[1], tz = t / (g.n[0] * g.n[1]);
    const int sx = pmod(tx - (e % 3 - 1), g.n[0]), sy = pmod(ty - ((e / 3) % 3 - 1), g.n[1]),
              sz = pmod(tz - (e / 9 - 1), g.n[2]);
    return ((int64_t)sz * g.n[1] + sy) * g.n[0] + sx;
}

// cnt[t] = number of particles arriving in t, counts[t] = the slots they take
// (cnt rounded up to a multiple of align); both 0 at t = ncell, for the scans.
__global__ void k_new_counts(const int* __restrict__ m, Geom g, int64_t ncell, int align, int* __restrict__ counts,
                             int* __restrict__ cnt)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t > ncell) return;
    int n = 0;
    if (t < ncell)
        for (int e = 0; e < 27; e++)
            if (valid_move(e, g)) n += m[source_of((int)t, e, g) * 27 + e];
    cnt[t] = n;
    counts[t] = (n + align - 1) / align * align;
}

// base[s][e] for every source s = t - e, in canonical move order.
__global__ void k_bases(const int* __restrict__ m, const int* __restrict__ new_start, Geom g, int64_t ncell,
                        int* __restrict__ base)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= ncell) return;
    int run = new_start[t];
    for (int e = 0; e < 27; e++) {
        if (!valid_move(e, g)) continue;
        const int64_t sidx = source_of((int)t, e, g) * 27 + e;
        base[sidx] = run;
        run += m[sidx];
    }
}

template <typename T> using Arrays = PArrays<T>;

// compute_dt over the cells (Evolve.cpp:40-56) from the combined fields and
// the auxiliaries.  HSI = N - conj(Nbar) + matter = M; F - Fbar keeps Re and the
// diagonal of Phi and takes Im(F - Fbar) = Im(Phi) - 2 Im(Fbar).  The
// auxiliaries hold S_ii = N_ii + Nbar_ii, so N_ii = (S_ii + M_ii - m00 d_i0) / 2
// and Nbar_ii = (S_ii - M_ii + m00 d_i0) / 2, then Im(Fbar_ab) per direction.
template <int NF>
__global__ void k_dt_reduce(const double* __restrict__ grid, const double* __restrict__ aux, int64_t ncell,
                            double cell_volume, double m00, int adaptive, unsigned long long* red)
{
    constexpr int nh = NH<NF>;
    constexpr int QG = 4 * nh;
    constexpr int NAUX = NF + 3 * (NF * (NF - 1) / 2);
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    double Vad = 0.0, Vst = 0.0;
    if (t < ncell) {
        const double* G = grid + t * QG;
        const double* A = aux + t * NAUX;
        // the adaptive bound (only where max_adaptive_speedup > 1 uses it: the
        // stage kernel then deposits Im(Fbar) as well)
        if (adaptive) {
            double h[nh];
#pragma unroll
            for (int c = 0; c < nh; c++) h[c] = G[c];
            double V2 = su_mag2<NF, double>(h);
#pragma unroll
            for (int d = 0; d < 3; d++) {
#pragma unroll
                for (int c = 0; c < nh; c++) h[c] = G[(d + 1) * nh + c];
                int o = NF;
#pragma unroll
                for (int x = 0; x < NF; x++)
#pragma unroll
                    for (int y = x + 1; y < NF; y++) {
                        h[herm_im<NF>(x, y)] -= 2.0 * A[o + d];
                        o += 3;
                    }
                V2 += su_mag2<NF, double>(h);
            }
            Vad = sqrt(V2) * M_SQRT2 * phys::GF / cell_volume;
        }
#pragma unroll
        for (int x = 0; x < NF; x++) {
            const double Mxx = G[herm_re<NF>(x, x)] - (x == 0 ? m00 : 0.0);
            const double Nxx = 0.5 * (A[x] + Mxx), Nbxx = 0.5 * (A[x] - Mxx);
            Vst = Vst > Nxx ? Vst : Nxx;
            Vst = Vst > Nbxx ? Vst : Nbxx;
        }
        Vst = Vst > m00 ? Vst : m00;
        Vst *= 2.0 * NF * M_SQRT2 * phys::GF / cell_volume;
    }
    for (int off = 32; off > 0; off >>= 1) {
        Vad = fmax(Vad, __shfl_xor(Vad, off));
        Vst = fmax(Vst, __shfl_xor(Vst, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&red[0], (unsigned long long)__double_as_longlong(Vad > 0 ? Vad : 0.0));
        atomicMax(&red[1], (unsigned long long)__double_as_longlong(Vst > 0 ? Vst : 0.0));
    }
}

struct DtParams {
    double dt_translation, flavor_cfl, max_adaptive_speedup, Vvac_max, fixed_dt;
    int use_flavor;
};

// main.cpp:193-198 and Evolve.cpp:57-88: the next dt from the reductions.
__global__ void k_dt_finalize(const unsigned long long* red, DtParams p, double* dt_next, unsigned long long* err)
{
    if (threadIdx.x != 0) return;
    if (p.fixed_dt > 0) { *dt_next = p.fixed_dt; return; }
    double dt_flavor = 0.0;
    if (p.use_flavor) {
        const double vad = __longlong_as_double((long long)red[0]) + p.Vvac_max;
        const double vst = __longlong_as_double((long long)red[1]) + p.Vvac_max;
        const double dfa = phys::hbar / vad * p.flavor_cfl;
        const double dfs = phys::hbar / vst * p.flavor_cfl;
        dt_flavor = dfs;
        if (p.max_adaptive_speedup > 1) dt_flavor = fmin(dfs * p.max_adaptive_speedup, dfa);
    }
    double dt = 0.0;
    if (p.dt_translation != 0.0 && dt_flavor != 0.0) dt = fmin(p.dt_translation, dt_flavor);
    else if (p.dt_translation != 0.0) dt = p.dt_translation;
    else if (dt_flavor != 0.0) dt = dt_flavor;
    else atomicMin(err, status_word_noindex(EMU_ERR_DT));
    *dt_next = dt;
}

// End of step: particle time advanced by the four weighted stages, dt <- next.
__global__ void k_step_end(double* time, double* dt, const double* dt_next, int64_t* steps)
{
#pragma clang fp contract(off)
    if (threadIdx.x != 0) return;
    const double h = *dt;
    double t = *time;
    t = h * (1. / 6.) * 1.0 + t;
    t = h * (1. / 3.) * 1.0 + t;
    t = h * (1. / 3.) * 1.0 + t;
    t = h * (1. / 6.) * 1.0 + t;
    *time = t;
    *dt = *dt_next;
    *steps += 1;
}

__global__ void k_fill_int(int* p, int64_t n, int v)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) p[t] = v;
}

__global__ void k_fill_u64v(unsigned long long* p, int n, unsigned long long v)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = v;
}

// Load: bin every particle of a caller's SoA view (PIdx order).
__global__ void k_bin_input(PView P, Geom g, int* cellidx, int* counts)
{
    for (int64_t ip = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; ip < P.np; ip += (int64_t)gridDim.x * blockDim.x) {
        int nc[3];
        for (int d = 0; d < 3; d++) nc[d] = bin1(P.pos[d][ip], g.plo[d], g.dxi[d], g.n[d]);
        const int c = (nc[2] * g.n[1] + nc[1]) * g.n[0] + nc[0];
        cellidx[ip] = c;
        atomicAdd(&counts[c], 1);
    }
}

__global__ void k_iota(int* p, int64_t n)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        p[t] = (int)t;
}

// Load: cnt[c] = particles binned in c; counts[c] becomes the slots they take
// (rounded up to a multiple of align).  counts[ncell] and cnt[ncell] stay 0.
__global__ void k_pad_counts(int* __restrict__ counts, int* __restrict__ cnt, int64_t ncell, int align)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t <= ncell; t += (int64_t)gridDim.x * blockDim.x) {
        const int n = t < ncell ? counts[t] : 0;
        cnt[t] = n;
        counts[t] = (n + align - 1) / align * align;
    }
}

// Load: the j-th particle in (cell, input index) order is input particle
// order[j] of cell key[j]; it takes slot start[key] + (j - ustart[key]), where
// ustart are the cells' first positions back to back and start their slots.
template <int NF, typename T>
__global__ void k_load(PView P, const int64_t* ids, const int* __restrict__ order, const int* __restrict__ key,
                       const int* __restrict__ start, const int* __restrict__ ustart, Arrays<T> dst, int64_t np,
                       int64_t stride, double nscale)
{
    using I = PIdx<NF>;
    constexpr int nh = NH<NF>;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < np; j += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ip = order[j];
        const int c = key[j];
        const int64_t k = start[c] + (j - ustart[c]);
        for (int d = 0; d < 3; d++) dst.pos[d][k] = P.pos[d][ip];
        for (int d = 0; d < 4; d++) dst.pup[d][k] = T(P.r[I::pupx + d][ip]);
        dst.Nn[0][k] = T(P.r[I::N][ip] * nscale);
        dst.Nn[1][k] = T(P.r[I::Nbar][ip] * nscale);
        dst.Ll[0][k] = T(P.r[I::L][ip]);
        dst.Ll[1][k] = T(P.r[I::Lbar][ip]);
        dst.id[k] = ids ? ids[ip] : ip;
        for (int c2 = 0; c2 < nh; c2++) {
            dst.f[(int64_t)c2 * stride + k] = T(P.r[I::f + c2][ip]);
            dst.f[(int64_t)(nh + c2) * stride + k] = T(P.r[I::fbar + c2][ip]);
        }
    }
}

// Store: one wave per cell copies its cnt[c] particles from slots start[c]..
// to the caller's positions ustart[c].. (cell order, no gaps).
template <int NF, typename T>
__global__ void k_store(Arrays<T> src, PView P, int64_t* ids, const double* time, const int* __restrict__ start,
                        const int* __restrict__ cnt, const int* __restrict__ ustart, int64_t ncell, int64_t stride,
                        double nscale)
{
    using I = PIdx<NF>;
    constexpr int nh = NH<NF>;
    const double t = *time;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t c = blockIdx.x * (int64_t)(blockDim.x / 64) + threadIdx.x / 64; c < ncell; c += waves) {
        const int n = cnt[c];
        for (int i = threadIdx.x % 64; i < n; i += 64) {
            const int64_t k = (int64_t)start[c] + i, ip = (int64_t)ustart[c] + i;
            for (int d = 0; d < 3; d++) {
                P.pos[d][ip] = src.pos[d][k];
                P.r[I::x + d][ip] = src.pos[d][k];
            }
            P.r[I::time][ip] = t;
            for (int d = 0; d < 4; d++) P.r[I::pupx + d][ip] = double(src.pup[d][k]);
            P.r[I::N][ip] = double(src.Nn[0][k]) / nscale;
            P.r[I::Nbar][ip] = double(src.Nn[1][k]) / nscale;
            P.r[I::L][ip] = double(src.Ll[0][k]);
            P.r[I::Lbar][ip] = double(src.Ll[1][k]);
            for (int c2 = 0; c2 < nh; c2++) {
                P.r[I::f + c2][ip] = double(src.f[(int64_t)c2 * stride + k]);
                P.r[I::fbar + c2][ip] = double(src.f[(int64_t)(nh + c2) * stride + k]);
            }
            if (ids) ids[ip] = src.id[k];
        }
    }
}

// max(|N|, |Nbar|) as the bit pattern of a non-negative double (those order
// like unsigned integers)
__global__ void k_absmax2(const double* __restrict__ a, const double* __restrict__ b, int64_t np,
                          unsigned long long* red)
{
    double v = 0.0;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < np; t += (int64_t)gridDim.x * blockDim.x)
        v = fmax(v, fmax(fabs(a[t]), fabs(b[t])));
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0) atomicMax(red, (unsigned long long)__double_as_longlong(v));
}

__global__ void k_min_pupt_t(const double* __restrict__ pupt, int64_t np, unsigned long long* red)
{
    double v = INFINITY;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < np; t += (int64_t)gridDim.x * blockDim.x)
        v = fmin(v, pupt[t]);
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0) atomicMin(red, (unsigned long long)__double_as_longlong(v));
}

// Diagnostic: per cell N_01 = sum_p w N f01 and Nbar_01 (Re, Im), deposited
// with the particle shape factors, then the cell sums of |N_01| and |Nbar_01|.
// One wave per source cell takes its particles 64 at a time; for each of the
// 27 target cells the four weighted values are summed across the wave and
// lane 0 adds them to the target (108 atomics per chunk, not per particle).
__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

template <int NF, typename T>
__global__ __launch_bounds__(64) void k_diag_deposit(Arrays<T> A, const int* __restrict__ cell_start,
                                                     const int* __restrict__ cell_cnt, Geom g, int64_t np, double* out4)
{
    const int cell = blockIdx.x;
    const int lane = threadIdx.x;
    const int beg = cell_start[cell], end = beg + cell_cnt[cell];
    for (int base = beg; base < end; base += 64) {
        const int ip = base + lane;
        const bool act = ip < end;
        const int ipl = act ? ip : beg;
        int idx[3];
        double wd[3][3];
        for (int d = 0; d < 3; d++) shape3((A.pos[d][ipl] - g.plo[d]) * g.dxi[d], g.order[d], idx[d], wd[d]);
        const double N = act ? double(A.Nn[0][ipl]) : 0.0, Nb = act ? double(A.Nn[1][ipl]) : 0.0;
        const double v[4] = {N * (double)A.f[(int64_t)herm_re<NF>(0, 1) * np + ipl],
                             N * (double)A.f[(int64_t)herm_im<NF>(0, 1) * np + ipl],
                             Nb * (double)A.f[(int64_t)(NH<NF> + herm_re<NF>(0, 1)) * np + ipl],
                             Nb * (double)A.f[(int64_t)(NH<NF> + herm_im<NF>(0, 1)) * np + ipl]};
        // every particle of the cell has the same first stencil cell per dimension
        // (the stencil is centred on the particle's own cell)
        const int cx = cell % g.n[0], cy = (cell / g.n[0]) % g.n[1], cz = cell / (g.n[0] * g.n[1]);
        for (int t = 0; t < 27; t++) {
            const int ii = t % 3, jj = (t / 3) % 3, kk = t / 9;
            const double w = wd[0][ii] * wd[1][jj] * wd[2][kk];
            double s4[4];
            for (int q = 0; q < 4; q++) s4[q] = wave_sum(w * v[q]);
            if (lane == 0 && (s4[0] != 0.0 || s4[1] != 0.0 || s4[2] != 0.0 || s4[3] != 0.0)) {
                const int64_t c = ((int64_t)pmod(cz + kk - 1, g.n[2]) * g.n[1] + pmod(cy + jj - 1, g.n[1])) * g.n[0] +
                                  pmod(cx + ii - 1, g.n[0]);
                for (int q = 0; q < 4; q++) unsafeAtomicAdd(&out4[c * 4 + q], s4[q]);
            }
        }
        (void)idx;
    }
}

__global__ void k_diag_reduce(const double* __restrict__ g4, int64_t ncell, double* sums)
{
    double a = 0, b = 0;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
        a += sqrt(g4[4 * c] * g4[4 * c] + g4[4 * c + 1] * g4[4 * c + 1]);
        b += sqrt(g4[4 * c + 2] * g4[4 * c + 2] + g4[4 * c + 3] * g4[4 * c + 3]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    if ((threadIdx.x & 63) == 0) {
        unsafeAtomicAdd(&sums[0], a);
        unsafeAtomicAdd(&sums[1], b);
    }
}

} // namespace
} // namespace emu

using namespace emu;

// ---------------------------------------------------------------------------
// the state object
// ---------------------------------------------------------------------------
struct emu_state {
    emu_ctx* ctx;
    int64_t np;
    int precision; // EMU_PLAIN_FP64 / EMU_FP32_BULK
    int nf, nh, QG, naux;
    int64_t ncell;
    size_t tsize;
    // Slots: cell c's particles sit in slots cell_start[c] .. cell_start[c] +
    // cell_cnt[c] - 1, and every cell's run starts at a multiple of `align`
    // (1: back to back; 128 bytes of T otherwise, so that a wave's loads and
    // stores of a cell's particles cover whole cache lines that no other cell
    // shares).  cap slots (the stride of the per-particle arrays) hold any
    // layout: np + ncell (align - 1).
    int align = 1;
    int64_t cap = 1;
    // persistent arrays, two copies for the sort
    void* block[2] = {nullptr, nullptr};
    // scratch
    void* Y = nullptr;
    void* acc = nullptr;
    int* newcell = nullptr;
    int* rank = nullptr;
    int* counts = nullptr;         // ncell + 1: the scan's input (padded counts)
    int* cell_start = nullptr;     // ncell + 1
    int* new_start = nullptr;      // ncell + 1
    int* cell_cnt = nullptr;       // ncell + 1: particles per cell (current order)
    int* new_cnt = nullptr;        // ncell + 1: the same after this step's moves
    int* mcount = nullptr;         // [ncell][27] particles per (old cell, move)
    int* mbase = nullptr;          // [ncell][27] destination of the first of them
    unsigned char* mv = nullptr;   // move index per particle
    double* grid[2] = {nullptr, nullptr};
    double* aux = nullptr;
    double* scal = nullptr;        // [0] time [1] dt [2] dt_next
    int64_t* steps = nullptr;
    unsigned long long* red = nullptr; // [0..1] dt maxima, [2] error, [3] pupt min
    void* cub_tmp = nullptr;
    size_t cub_bytes = 0;
    int cur = 0;      // which persistent copy is current
    int gcur = 0;     // grid holding the deposit of the current state
    bool evolved = false; // a step was taken since the last load
    double Vvac_max = 0.0;
    double m00 = 0.0;
    // The particle weights are held as N * nscale, nscale = 2^-k chosen at load
    // so that the largest weight is in [1, 2): FP32 cannot hold N = n dV for
    // large cells (1e54 for a (1e7 cm)^3 cell), and a power of two scales every
    // product and sum exactly.  The deposited fields carry the same factor,
    // which coef, the matter term and the dt estimate take back out.  The FP64
    // state keeps nscale = 1.
    double nscale = 1.0;
    // timing: an event pair around every profiled launch, by kernel class
    int profile = 0;
    struct Mark { int k; hipEvent_t a, b; };
    std::vector<Mark> marks;
    // the sort plan runs on a stream of its own, beside stages 2 and 3 (it
    // needs stage 1's end-of-step cells, and only stage 4 reads it), and so
    // does compute_dt, beside stage 4 and the step-end deposit
    hipStream_t side = nullptr;
    hipEvent_t ev_stage1 = nullptr, ev_plan = nullptr, ev_stage3 = nullptr, ev_dt = nullptr;
};

namespace {

template <typename T> Arrays<T> arrays_of(emu_state* st, int copy)
{
    Arrays<T> A;
    char* b = (char*)st->block[copy];
    const int64_t np = st->cap;
    for (int d = 0; d < 3; d++) { A.pos[d] = (double*)b; b += np * sizeof(double); }
    A.id = (int64_t*)b;
    b += np * sizeof(int64_t);
    for (int d = 0; d < 4; d++) { A.pup[d] = (T*)b; b += np * sizeof(T); }
    for (int d = 0; d < 2; d++) { A.Nn[d] = (T*)b; b += np * sizeof(T); }
    for (int d = 0; d < 2; d++) { A.Ll[d] = (T*)b; b += np * sizeof(T); }
    A.f = (T*)b;
    return A;
}

// The ids of the current copy (placed after the positions in either precision).
const int64_t* current_ids(emu_state* st) { return arrays_of<double>(st, st->cur).id; }

size_t persistent_bytes(int64_t np, int nh, size_t tsize)
{
    return (size_t)np * (3 * sizeof(double) + sizeof(int64_t) + (8 + 2 * nh) * tsize);
}

unsigned grid_blocks(int64_t n, int bs = 256)
{
    return (unsigned)std::min<int64_t>(65535 * 16, std::max<int64_t>(1, (n + bs - 1) / bs));
}

int ensure_cub(emu_state* st, size_t need)
{
    if (need > st->cub_bytes) {
        if (st->cub_tmp) (void)hipFree(st->cub_tmp);
        st->cub_tmp = nullptr;
        st->cub_bytes = 0;
        if (hipMalloc(&st->cub_tmp, need) != hipSuccess) return EMU_ERR_NOMEM;
        st->cub_bytes = need;
    }
    return EMU_OK;
}

// out = exclusive scan of in (ncell + 1 entries, in[ncell] = 0), on stream s
// (the context's when null)
int scan_counts(emu_state* st, const int* in, int* out, hipStream_t s = nullptr)
{
    if (!s) s = st->ctx->stream;
    size_t need = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, need, in, out, (int)(st->ncell + 1), s);
    if (ensure_cub(st, need)) return set_error(st->ctx, EMU_ERR_NOMEM, "scan scratch");
    EMU_HIP(st->ctx, hipcub::DeviceScan::ExclusiveSum(st->cub_tmp, st->cub_bytes, in, out, (int)(st->ncell + 1), s));
    return EMU_OK;
}


template <int NF, typename T, int WX, int WY, int WZ, int S, int AUX>
void launch_stage_k(emu_state* st, StageArgs<T>& a)
{
    constexpr int MODE = stage_mode<T, S, AUX>();
    constexpr bool PM = MODE >= 1;
    constexpr int WPB = stage_waves_per_block<MODE, AUX>(); // waves per workgroup
    using PT = Patch<WX, WY, WZ, WPB, PM, stage_cells_per_wave<T, WX, MODE>()>;
    const Geom& g = st->ctx->g;
    const int64_t npatch = (int64_t)((g.n[0] + PT::PX - 1) / PT::PX) * ((g.n[1] + PT::PY - 1) / PT::PY) *
                           ((g.n[2] + PT::PZ - 1) / PT::PZ);
    const int64_t nblk = PM ? npatch : (npatch + WPB - 1) / WPB;
    a.blocks_per_xcd = (int)((nblk + 7) / 8);
    const size_t lds = (size_t)stage_lds_doubles_per_block<NF, T, WX, WY, WZ, WPB, S, AUX, MODE>() * sizeof(double);
    hipLaunchKernelGGL((k_stage<NF, T, WX, WY, WZ, WPB, S, AUX, MODE>), dim3((unsigned)(8 * a.blocks_per_xcd)),
                       dim3(64 * WPB), lds, st->ctx->stream, a);
}

template <int NF, typename T, int WX, int WY, int WZ>
void launch_stage_t(emu_state* st, StageArgs<T>& a)
{
    // the compute_dt auxiliaries ride on the deposit that compute_dt reads
    const int aux = a.stage == 3 || a.stage == 0 ? a.want_aux : 0;
    switch (a.stage) {
    case 0:
        if (aux == 2) launch_stage_k<NF, T, WX, WY, WZ, 0, 2>(st, a);
        else if (aux == 1) launch_stage_k<NF, T, WX, WY, WZ, 0, 1>(st, a);
        else launch_stage_k<NF, T, WX, WY, WZ, 0, 0>(st, a);
        break;
    case 1: launch_stage_k<NF, T, WX, WY, WZ, 1, 0>(st, a); break;
    case 2: launch_stage_k<NF, T, WX, WY, WZ, 2, 0>(st, a); break;
    case 3:
        if (aux == 2) launch_stage_k<NF, T, WX, WY, WZ, 3, 2>(st, a);
        else if (aux == 1) launch_stage_k<NF, T, WX, WY, WZ, 3, 1>(st, a);
        else launch_stage_k<NF, T, WX, WY, WZ, 3, 0>(st, a);
        break;
    default: launch_stage_k<NF, T, WX, WY, WZ, 4, 0>(st, a); break;
    }
}

// Stencil shapes: 3 cells per dimension with shape order 1 or 2, 1 cell with
// order 0 or in a 1-cell dimension.  (3,3,3), (1,1,3) and (1,1,1) have their
// own kernels; other shapes run the (3,3,3) kernel, whose extra entries carry
// zero weight.
template <int NF, typename T> void launch_stage_shape(emu_state* st, StageArgs<T>& a)
{
    const int* o = st->ctx->g.order;
    const int wx = o[0] > 0 ? 3 : 1, wy = o[1] > 0 ? 3 : 1, wz = o[2] > 0 ? 3 : 1;
    if (wx == 1 && wy == 1 && wz == 1) launch_stage_t<NF, T, 1, 1, 1>(st, a);
    else if (wx == 1 && wy == 1) launch_stage_t<NF, T, 1, 1, 3>(st, a);
    else launch_stage_t<NF, T, 3, 3, 3>(st, a);
}

template <typename T>
StageArgs<T> stage_args(emu_state* st, int stage, const double* gin, double* gout, bool want_aux)
{
    StageArgs<T> a;
    a.cur = arrays_of<T>(st, st->cur);
    a.dst = arrays_of<T>(st, 1 - st->cur);
    a.Y = (T*)st->Y;
    a.acc = (T*)st->acc;
    a.newcell = st->newcell;
    a.mbase = st->mbase;
    a.rank = st->rank;
    a.mv = st->mv;
    a.cell_start = st->cell_start;
    a.cell_cnt = st->cell_cnt;
    a.gin = gin;
    a.gout = gout;
    a.aux = st->aux;
    a.dt = st->scal + 1;
    a.err = st->red + 2;
    a.np = st->cap;
    a.ncell = st->ncell;
    a.stage = stage;
    // the diagonal sums for the conservative flavour bound; Im(Fbar) as well
    // where the adaptive bound is used (max_adaptive_speedup > 1).  FP64 deposits
    // the full set either way: measured at 64^3 x 92, its auxiliary stage ran
    // 2% slower with the diagonal sums alone (FP32-bulk's 1% faster).
    a.want_aux = want_aux ? (st->ctx->p.max_adaptive_speedup > 1 || sizeof(T) == 8 ? 2 : 1) : 0;
    a.blocks_per_xcd = 0;
    a.g = st->ctx->g;
    const emu_ctx* ctx = st->ctx;
    for (int c = 0; c < 9; c++) a.vac[c] = c < st->nh ? ctx->M2[c] * phys::c4 / 2.0 / phys::hbar : 0.0;
    a.coef = M_SQRT2 * phys::GF * ctx->g.inv_cell_volume / phys::hbar / st->nscale;
    a.maxError = ctx->p.maxError;
    return a;
}

void launch_stage(emu_state* st, int stage, const double* gin, double* gout, bool want_aux)
{
    if (st->precision == EMU_FP32_BULK) {
        auto a = stage_args<float>(st, stage, gin, gout, want_aux);
        if (st->nf == 2) launch_stage_shape<2, float>(st, a);
        else launch_stage_shape<3, float>(st, a);
    } else {
        auto a = stage_args<double>(st, stage, gin, gout, want_aux);
        if (st->nf == 2) launch_stage_shape<2, double>(st, a);
        else launch_stage_shape<3, double>(st, a);
    }
}

void fill_grid(emu_state* st, double* g, bool with_aux)
{
    hipLaunchKernelGGL(k_fill_grid, dim3(grid_blocks(st->ncell * st->QG)), dim3(256), 0, st->ctx->stream, g,
                       st->ncell, st->QG, st->m00 * st->nscale, with_aux ? st->aux : nullptr, st->naux);
}

int dt_from_grid(emu_state* st, const double* g, double fixed_dt, hipStream_t s = nullptr)
{
    emu_ctx* ctx = st->ctx;
    const emu_params& p = ctx->p;
    if (!s) s = ctx->stream;
    hipLaunchKernelGGL(k_fill_u64v, dim3(1), dim3(64), 0, s, st->red, 2, 0ull);
    const bool use_flavor = p.flavor_cfl_factor > 0.0;
    const int adaptive = p.max_adaptive_speedup > 1 ? 1 : 0;
    if (use_flavor && fixed_dt <= 0) {
        if (st->nf == 2)
            hipLaunchKernelGGL(k_dt_reduce<2>, dim3(grid_blocks(st->ncell)), dim3(256), 0, s, g, st->aux,
                               st->ncell, ctx->g.cell_volume * st->nscale, st->m00 * st->nscale, adaptive, st->red);
        else
            hipLaunchKernelGGL(k_dt_reduce<3>, dim3(grid_blocks(st->ncell)), dim3(256), 0, s, g, st->aux,
                               st->ncell, ctx->g.cell_volume * st->nscale, st->m00 * st->nscale, adaptive, st->red);
    }
    DtParams dp;
    dp.dt_translation = p.cfl_factor > 0.0
                            ? std::min(std::min(ctx->g.dx[0], ctx->g.dx[1]), ctx->g.dx[2]) / phys::c * p.cfl_factor
                            : 0.0;
    dp.flavor_cfl = p.flavor_cfl_factor;
    dp.max_adaptive_speedup = p.max_adaptive_speedup;
    dp.Vvac_max = st->Vvac_max;
    dp.fixed_dt = fixed_dt;
    dp.use_flavor = use_flavor;
    hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(64), 0, s, st->red, dp, st->scal + 2, st->red + 2);
    EMU_LAUNCH_CHECK(ctx, "compute_dt");
    return EMU_OK;
}

// Kernel classes: 0 stage kernels, 1 grid fills, 2 sort, 3 dt reduction.
void ev_record(emu_state* st, int k, bool start, hipStream_t stream = nullptr)
{
    if (!st->profile) return;
    hipEvent_t e;
    (void)hipEventCreate(&e);
    (void)hipEventRecord(e, stream ? stream : st->ctx->stream);
    if (start) st->marks.push_back({k, e, nullptr});
    else st->marks.back().b = e;
}

} // namespace

extern "C" int emu_state_destroy(emu_state* st);

extern "C" int emu_state_create(emu_ctx* ctx, int64_t np, int precision, emu_state** out)
{
    const int layout = precision & ~0xff;
    precision &= 0xff;
    if (!ctx || !out || np < 0 || (precision != EMU_PLAIN_FP64 && precision != EMU_FP32_BULK) ||
        (layout != 0 && layout != EMU_STATE_SLOTS_PACKED && layout != EMU_STATE_SLOTS_ALIGNED))
        return EMU_ERR_ARG;
    if (np > (int64_t)INT32_MAX) return set_error(ctx, EMU_ERR_UNSUPPORTED, "more than 2^31-1 particles per state");
    *out = nullptr;
    auto* st = new emu_state();
    st->ctx = ctx;
    st->np = np;
    st->precision = precision;
    st->nf = ctx->p.nf;
    st->nh = st->nf * st->nf;
    st->QG = 4 * st->nh;
    st->naux = st->nf + 3 * (st->nf * (st->nf - 1) / 2);
    st->ncell = (int64_t)ctx->g.n[0] * ctx->g.n[1] * ctx->g.n[2];
    st->tsize = precision == EMU_FP32_BULK ? sizeof(float) : sizeof(double);
    st->m00 = ctx->p.rho * ctx->p.Ye / phys::Mp / ctx->g.inv_cell_volume;
    // slot alignment: 128 bytes of T; by default only where the cells hold on
    // average two alignments' worth of particles (padding <= np / 2)
    const int a128 = (int)(128 / st->tsize);
    bool aligned = layout == EMU_STATE_SLOTS_ALIGNED || (layout == 0 && np >= 2 * a128 * st->ncell);
    if (np == 0 || np + st->ncell * (a128 - 1) > (int64_t)INT32_MAX) aligned = false;
    st->align = aligned ? a128 : 1;
    st->cap = std::max<int64_t>(np + st->ncell * (st->align - 1), 1);
    const int64_t npa = st->cap;
    const size_t pb = persistent_bytes(npa, st->nh, st->tsize);
    bool ok = hipMalloc(&st->block[0], pb) == hipSuccess && hipMalloc(&st->block[1], pb) == hipSuccess &&
              hipMalloc(&st->Y, (size_t)npa * 2 * st->nh * st->tsize) == hipSuccess &&
              hipMalloc(&st->acc, (size_t)npa * 2 * st->nh * st->tsize) == hipSuccess &&
              hipMalloc(&st->newcell, (size_t)npa * sizeof(int)) == hipSuccess &&
              hipMalloc(&st->rank, (size_t)npa * sizeof(int)) == hipSuccess &&
              hipMalloc(&st->counts, (size_t)(st->ncell + 1) * sizeof(int)) == hipSuccess &&
              hipMalloc(&st->cell_start, (size_t)(st->ncell + 1) * sizeof(int)) == hipSuccess &&
              hipMalloc(&st->new_start, (size_t)(st->ncell + 1) * sizeof(int)) == hipSuccess &&
              hipMalloc(&st->cell_cnt, (size_t)(st->ncell + 1) * sizeof(int)) == hipSuccess &&
              hipMalloc(&st->new_cnt, (size_t)(st->ncell + 1) * sizeof(int)) == hipSuccess &&
              hipMalloc(&st->mcount, (size_t)st->ncell * 27 * sizeof(int)) == hipSuccess &&
              hipMalloc(&st->mbase, (size_t)st->ncell * 27 * sizeof(int)) == hipSuccess &&
              hipMalloc(&st->mv, (size_t)npa) == hipSuccess &&
              hipMalloc(&st->grid[0], (size_t)st->ncell * st->QG * sizeof(double)) == hipSuccess &&
              hipMalloc(&st->grid[1], (size_t)st->ncell * st->QG * sizeof(double)) == hipSuccess &&
              hipMalloc(&st->aux, (size_t)st->ncell * st->naux * sizeof(double)) == hipSuccess &&
              hipMalloc(&st->scal, 4 * sizeof(double)) == hipSuccess &&
              hipMalloc(&st->steps, sizeof(int64_t)) == hipSuccess &&
              hipMalloc(&st->red, 8 * sizeof(unsigned long long)) == hipSuccess;
    // default priority, and no implicit synchronisation with the legacy stream
    ok = ok && hipStreamCreateWithFlags(&st->side, hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&st->ev_stage1, hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&st->ev_plan, hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&st->ev_stage3, hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&st->ev_dt, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        emu_state_destroy(st);
        return set_error(ctx, EMU_ERR_NOMEM, "state allocation failed");
    }
    *out = st;
    return EMU_OK;
}

extern "C" int emu_state_destroy(emu_state* st)
{
    if (!st) return EMU_OK;
    void* ptrs[] = {st->block[0], st->block[1], st->Y,        st->acc,     st->newcell, st->rank,
                    st->counts,   st->cell_start, st->new_start, st->cell_cnt, st->new_cnt, st->mcount,
                    st->mbase,    st->mv,         st->grid[0], st->grid[1], st->aux,
                    st->scal,     st->steps,      st->red,     st->cub_tmp};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (auto& m : st->marks) {
        (void)hipEventDestroy(m.a);
        if (m.b) (void)hipEventDestroy(m.b);
    }
    if (st->side) {
        (void)hipStreamSynchronize(st->side);
        (void)hipStreamDestroy(st->side);
    }
    for (hipEvent_t ev : {st->ev_stage1, st->ev_plan, st->ev_stage3, st->ev_dt})
        if (ev) (void)hipEventDestroy(ev);
    delete st;
    return EMU_OK;
}

template <typename T> static int load_t(emu_state* st, const PView& P, const int64_t* ids)
{
    emu_ctx* ctx = st->ctx;
    hipStream_t s = ctx->stream;
    const int64_t np = st->np;
    hipLaunchKernelGGL(k_fill_int, dim3(grid_blocks(st->ncell + 1)), dim3(256), 0, s, st->counts, st->ncell + 1, 0);
    if (np > 0)
        hipLaunchKernelGGL(k_bin_input, dim3(grid_blocks(np)), dim3(256), 0, s, P, ctx->g, st->newcell, st->counts);
    // the cells' first positions back to back (new_start, scratch here), then
    // their padded slot runs (cell_start)
    int e = scan_counts(st, st->counts, st->new_start);
    if (e) return e;
    hipLaunchKernelGGL(k_pad_counts, dim3(grid_blocks(st->ncell + 1)), dim3(256), 0, s, st->counts, st->cell_cnt,
                       st->ncell, st->align);
    e = scan_counts(st, st->counts, st->cell_start);
    if (e) return e;
    if (np == 0) return EMU_OK;
    st->nscale = 1.0;
    if (st->precision == EMU_FP32_BULK) {
        using I2 = PIdx<2>;
        const int iN = I2::N, iNb = st->nf == 2 ? PIdx<2>::Nbar : PIdx<3>::Nbar;
        hipLaunchKernelGGL(k_fill_u64v, dim3(1), dim3(64), 0, s, st->red + 4, 1, 0ull);
        hipLaunchKernelGGL(k_absmax2, dim3(std::min(grid_blocks(np), 1024u)), dim3(256), 0, s, P.r[iN], P.r[iNb], np, st->red + 4);
        unsigned long long h = 0;
        EMU_HIP(ctx, hipMemcpyAsync(&h, st->red + 4, 8, hipMemcpyDeviceToHost, s));
        EMU_HIP(ctx, hipStreamSynchronize(s));
        double nmax;
        std::memcpy(&nmax, &h, 8);
        if (nmax > 0.0 && std::isfinite(nmax)) st->nscale = std::ldexp(1.0, -std::ilogb(nmax));
    }
    // stable radix sort of (cell, input index): sorted slot j <- input order[j]
    int *keys_out = nullptr, *order = nullptr;
    if (hipMalloc(&keys_out, np * sizeof(int)) != hipSuccess || hipMalloc(&order, np * sizeof(int)) != hipSuccess) {
        if (keys_out) (void)hipFree(keys_out);
        return set_error(ctx, EMU_ERR_NOMEM, "load sort scratch");
    }
    hipLaunchKernelGGL(k_iota, dim3(grid_blocks(np)), dim3(256), 0, s, st->rank, np);
    int end_bit = 1;
    while ((1ll << end_bit) < st->ncell) end_bit++;
    size_t need = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, need, st->newcell, keys_out, st->rank, order, (int)np, 0, end_bit,
                                             s);
    if (ensure_cub(st, need)) e = set_error(ctx, EMU_ERR_NOMEM, "radix sort scratch");
    if (!e && hipcub::DeviceRadixSort::SortPairs(st->cub_tmp, st->cub_bytes, st->newcell, keys_out, st->rank, order,
                                                 (int)np, 0, end_bit, s) != hipSuccess)
        e = set_error(ctx, EMU_ERR_HIP, "radix sort");
    if (!e) {
        Arrays<T> A = arrays_of<T>(st, st->cur);
        if (st->nf == 2)
            hipLaunchKernelGGL((k_load<2, T>), dim3(grid_blocks(np)), dim3(256), 0, s, P, ids, order, keys_out,
                               st->cell_start, st->new_start, A, np, st->cap, st->nscale);
        else
            hipLaunchKernelGGL((k_load<3, T>), dim3(grid_blocks(np)), dim3(256), 0, s, P, ids, order, keys_out,
                               st->cell_start, st->new_start, A, np, st->cap, st->nscale);
        if (hipGetLastError() != hipSuccess) e = set_error(ctx, EMU_ERR_HIP, "state_load");
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(keys_out);
    (void)hipFree(order);
    return e;
}

extern "C" int emu_state_load(emu_state* st, const emu_particles* parts, const int64_t* ids)
{
    if (!st || !parts || parts->np != st->np) return EMU_ERR_ARG;
    emu_ctx* ctx = st->ctx;
    PView P;
    P.np = parts->np;
    const int na = 12 + 2 * st->nh;
    for (int d = 0; d < 3; d++) P.pos[d] = parts->pos[d];
    for (int k = 0; k < EMU_MAX_ATTRIBS; k++) P.r[k] = k < na ? parts->rdata[k] : nullptr;
    int e = st->precision == EMU_FP32_BULK ? load_t<float>(st, P, ids) : load_t<double>(st, P, ids);
    if (e) return e;
    hipStream_t s = ctx->stream;
    // time of the particles (all particles share it; main.cpp:101-110)
    double t0 = 0.0;
    if (st->np > 0) EMU_HIP(ctx, hipMemcpyAsync(&t0, parts->rdata[PIdx<2>::time], sizeof(double), hipMemcpyDeviceToHost, s));
    // Vvac_max from the minimum energy (FlavoredNeutrinoContainerInit.cpp:657-660)
    hipLaunchKernelGGL(k_fill_u64v, dim3(1), dim3(64), 0, s, st->red, 4, ~0ull >> 1);
    if (st->np > 0)
        hipLaunchKernelGGL(k_min_pupt_t, dim3(std::min(grid_blocks(st->np), 1024u)), dim3(256), 0, s, parts->rdata[PIdx<2>::pupt], st->np,
                           st->red + 3);
    unsigned long long h;
    EMU_HIP(ctx, hipMemcpyAsync(&h, st->red + 3, 8, hipMemcpyDeviceToHost, s));
    EMU_HIP(ctx, hipStreamSynchronize(s));
    double pupt_min;
    std::memcpy(&pupt_min, &h, 8);
    double d[9] = {0};
    if (st->nf == 2) {
        for (int i = 0; i < 2; i++) d[herm_re<2>(i, i)] = ctx->p.mass[i] * ctx->p.mass[i];
        st->Vvac_max = std::sqrt(su_mag2<2, double>(d)) * phys::c4 / pupt_min;
    } else {
        for (int i = 0; i < 3; i++) d[herm_re<3>(i, i)] = ctx->p.mass[i] * ctx->p.mass[i];
        st->Vvac_max = std::sqrt(su_mag2<3, double>(d)) * phys::c4 / pupt_min;
    }
    double sc[4] = {t0, 0.0, 0.0, 0.0};
    EMU_HIP(ctx, hipMemcpyAsync(st->scal, sc, sizeof sc, hipMemcpyHostToDevice, s));
    EMU_HIP(ctx, hipMemsetAsync(st->steps, 0, sizeof(int64_t), s));
    hipLaunchKernelGGL(k_fill_u64v, dim3(1), dim3(64), 0, s, st->red + 2, 1, ~0ull);
    // initial deposit (main.cpp:118) and starting dt (main.cpp:206)
    st->gcur = 0;
    fill_grid(st, st->grid[0], true);
    launch_stage(st, 0, nullptr, st->grid[0], true);
    e = dt_from_grid(st, st->grid[0], 0.0);
    if (e) return e;
    EMU_HIP(ctx, hipMemcpyAsync(st->scal + 1, st->scal + 2, sizeof(double), hipMemcpyDeviceToDevice, s));
    EMU_HIP(ctx, hipStreamSynchronize(s));
    st->evolved = false;
    return EMU_OK;
}

extern "C" int emu_state_set_vvac_max(emu_state* st, double vvac_max)
{
    if (!st || !(vvac_max >= 0.0) || !std::isfinite(vvac_max)) return EMU_ERR_ARG;
    emu_ctx* ctx = st->ctx;
    if (st->evolved) return set_error(ctx, EMU_ERR_ARG, "Vvac_max is set between emu_state_load and the first step");
    st->Vvac_max = vvac_max;
    // grid[0] (and aux) still hold the load's deposit: the starting dt again
    // (main.cpp:206) with the new value
    const int e = dt_from_grid(st, st->grid[0], 0.0);
    if (e) return e;
    hipStream_t s = ctx->stream;
    EMU_HIP(ctx, hipMemcpyAsync(st->scal + 1, st->scal + 2, sizeof(double), hipMemcpyDeviceToDevice, s));
    EMU_HIP(ctx, hipStreamSynchronize(s));
    return EMU_OK;
}

extern "C" int emu_state_store(emu_state* st, emu_particles* parts, int64_t* ids)
{
    if (!st || !parts || parts->np != st->np) return EMU_ERR_ARG;
    emu_ctx* ctx = st->ctx;
    PView P;
    P.np = parts->np;
    const int na = 12 + 2 * st->nh;
    for (int d = 0; d < 3; d++) P.pos[d] = parts->pos[d];
    for (int k = 0; k < EMU_MAX_ATTRIBS; k++) P.r[k] = k < na ? parts->rdata[k] : nullptr;
    if (st->np > 0) {
        // the caller's positions of each cell's first particle (new_start is
        // scratch between steps)
        int e = scan_counts(st, st->cell_cnt, st->new_start);
        if (e) return e;
        const dim3 grid(grid_blocks(st->ncell, 4)), block(256); // one wave per cell
        const int* cs = st->cell_start;
        const int* cn = st->cell_cnt;
        const int* us = st->new_start;
        if (st->precision == EMU_FP32_BULK) {
            Arrays<float> A = arrays_of<float>(st, st->cur);
            if (st->nf == 2)
                hipLaunchKernelGGL((k_store<2, float>), grid, block, 0, ctx->stream, A, P, ids, st->scal, cs, cn, us,
                                   st->ncell, st->cap, st->nscale);
            else
                hipLaunchKernelGGL((k_store<3, float>), grid, block, 0, ctx->stream, A, P, ids, st->scal, cs, cn, us,
                                   st->ncell, st->cap, st->nscale);
        } else {
            Arrays<double> A = arrays_of<double>(st, st->cur);
            if (st->nf == 2)
                hipLaunchKernelGGL((k_store<2, double>), grid, block, 0, ctx->stream, A, P, ids, st->scal, cs, cn, us,
                                   st->ncell, st->cap, st->nscale);
            else
                hipLaunchKernelGGL((k_store<3, double>), grid, block, 0, ctx->stream, A, P, ids, st->scal, cs, cn, us,
                                   st->ncell, st->cap, st->nscale);
        }
    }
    EMU_LAUNCH_CHECK(ctx, "state_store");
    EMU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return EMU_OK;
}

extern "C" int emu_state_evolve(emu_state* st, int nsteps, double fixed_dt)
{
    if (!st || nsteps < 0) return EMU_ERR_ARG;
    emu_ctx* ctx = st->ctx;
    hipStream_t s = ctx->stream;
    const Geom& g = ctx->g;
    hipLaunchKernelGGL(k_fill_int, dim3(grid_blocks(st->ncell + 1)), dim3(256), 0, s, st->counts, st->ncell + 1, 0);
    // the deterministic sort assumes moves of at most one cell per step
    if (ctx->p.cfl_factor <= 0.0 || ctx->p.cfl_factor >= 1.0)
        if (fixed_dt <= 0 || fixed_dt * phys::c >= std::min(std::min(g.dx[0], g.dx[1]), g.dx[2]))
            return set_error(ctx, EMU_ERR_MOVE, "fused step needs c*dt < dx (0 < cfl_factor < 1)");
    if (fixed_dt > 0) EMU_HIP(ctx, hipMemcpyAsync(st->scal + 1, &fixed_dt, sizeof(double), hipMemcpyHostToDevice, s));
    if (nsteps > 0) st->evolved = true;
    for (int step = 0; step < nsteps; step++) {
        double* A = st->grid[st->gcur];     // holds the deposit of y_n
        double* B = st->grid[1 - st->gcur];
        const bool want_aux = ctx->p.flavor_cfl_factor > 0.0 && fixed_dt <= 0;
        ev_record(st, 1, true);
        fill_grid(st, B, false);
        ev_record(st, 1, false);
        ev_record(st, 0, true);
        launch_stage(st, 1, A, B, false);   // also bins the end-of-step positions
        ev_record(st, 0, false);
        // stable counting-sort plan by the new cell, on the side stream while
        // stages 2 and 3 run (it reads stage 1's cells and the current run
        // table; it writes only the plan, which stage 4 waits for)
        hipStream_t ss = st->side;
        EMU_HIP(ctx, hipEventRecord(st->ev_stage1, s));
        EMU_HIP(ctx, hipStreamWaitEvent(ss, st->ev_stage1, 0));
        ev_record(st, 2, true, ss);
        hipLaunchKernelGGL(k_move_ranks, dim3((unsigned)st->ncell), dim3(64), 0, ss, st->newcell, st->cell_start,
                           st->cell_cnt, current_ids(st), ctx->g, st->rank, st->mv, st->mcount, st->red + 2);
        hipLaunchKernelGGL(k_new_counts, dim3(grid_blocks(st->ncell + 1)), dim3(256), 0, ss, st->mcount, ctx->g,
                           st->ncell, st->align, st->counts, st->new_cnt);
        int e = scan_counts(st, st->counts, st->new_start, ss);
        if (e) return e;
        hipLaunchKernelGGL(k_bases, dim3(grid_blocks(st->ncell)), dim3(256), 0, ss, st->mcount, st->new_start, ctx->g,
                           st->ncell, st->mbase);
        ev_record(st, 2, false, ss);
        EMU_HIP(ctx, hipEventRecord(st->ev_plan, ss));
        ev_record(st, 1, true);
        fill_grid(st, A, false);
        ev_record(st, 1, false);
        ev_record(st, 0, true);
        launch_stage(st, 2, B, A, false);
        ev_record(st, 0, false);
        ev_record(st, 1, true);
        fill_grid(st, B, want_aux);
        ev_record(st, 1, false);
        ev_record(st, 0, true);
        launch_stage(st, 3, A, B, want_aux);
        ev_record(st, 0, false);
        // compute_dt from stage 3's deposit, also on the side stream: it reads B
        // beside stage 4 and writes the next dt, which k_step_end waits for
        EMU_HIP(ctx, hipEventRecord(st->ev_stage3, s));
        EMU_HIP(ctx, hipStreamWaitEvent(ss, st->ev_stage3, 0));
        ev_record(st, 3, true, ss);
        e = dt_from_grid(st, B, fixed_dt, ss);
        ev_record(st, 3, false, ss);
        if (e) return e;
        EMU_HIP(ctx, hipEventRecord(st->ev_dt, ss));
        EMU_HIP(ctx, hipStreamWaitEvent(s, st->ev_plan, 0));
        ev_record(st, 0, true);
        launch_stage(st, 4, B, nullptr, false); // writes y_{n+1} into the other copy, new order
        ev_record(st, 0, false);
        st->cur = 1 - st->cur;
        std::swap(st->cell_start, st->new_start);
        std::swap(st->cell_cnt, st->new_cnt);
        // deposit of y_{n+1} for the next step's stage 1
        ev_record(st, 1, true);
        fill_grid(st, A, false);
        ev_record(st, 1, false);
        ev_record(st, 0, true);
        launch_stage(st, 0, nullptr, A, false);
        ev_record(st, 0, false);
        EMU_HIP(ctx, hipStreamWaitEvent(s, st->ev_dt, 0));
        hipLaunchKernelGGL(k_step_end, dim3(1), dim3(64), 0, s, st->scal, st->scal + 1, st->scal + 2, st->steps);
        if (fixed_dt > 0) EMU_HIP(ctx, hipMemcpyAsync(st->scal + 1, &fixed_dt, sizeof(double), hipMemcpyHostToDevice, s));
        EMU_LAUNCH_CHECK(ctx, "state_evolve");
    }
    return EMU_OK;
}

extern "C" int emu_state_info(emu_state* st, double* time, double* dt, int64_t* steps, int64_t* bad_index)
{
    if (!st) return EMU_ERR_ARG;
    emu_ctx* ctx = st->ctx;
    double sc[4];
    int64_t n = 0;
    unsigned long long e = 0;
    EMU_HIP(ctx, hipMemcpyAsync(sc, st->scal, sizeof sc, hipMemcpyDeviceToHost, ctx->stream));
    EMU_HIP(ctx, hipMemcpyAsync(&n, st->steps, sizeof n, hipMemcpyDeviceToHost, ctx->stream));
    EMU_HIP(ctx, hipMemcpyAsync(&e, st->red + 2, sizeof e, hipMemcpyDeviceToHost, ctx->stream));
    EMU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (time) *time = sc[0];
    if (dt) *dt = sc[1];
    if (steps) *steps = n;
    return status_decode(e, bad_index);
}


extern "C" int emu_state_kernel_ms(emu_state* st, int k, double* avg_ms, int64_t* launches)
{
    if (!st || k < 0 || k > 3) return EMU_ERR_ARG;
    emu_ctx* ctx = st->ctx;
    EMU_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double tot = 0.0;
    int64_t n = 0;
    for (auto& m : st->marks) {
        if (m.k != k || !m.b) continue;
        float ms = 0;
        if (hipEventElapsedTime(&ms, m.a, m.b) == hipSuccess) {
            tot += ms;
            n++;
        }
    }
    if (avg_ms) *avg_ms = n ? tot / n : 0.0;
    if (launches) *launches = n;
    return EMU_OK;
}

extern "C" int emu_state_profile(emu_state* st, int enable)
{
    if (!st) return EMU_ERR_ARG;
    (void)hipStreamSynchronize(st->ctx->stream);
    for (auto& m : st->marks) {
        (void)hipEventDestroy(m.a);
        if (m.b) (void)hipEventDestroy(m.b);
    }
    st->marks.clear();
    st->profile = enable;
    return EMU_OK;
}

template <typename T> static void diag_t(emu_state* st, double* g4)
{
    Arrays<T> A = arrays_of<T>(st, st->cur);
    if (st->nf == 2)
        hipLaunchKernelGGL((k_diag_deposit<2, T>), dim3((unsigned)st->ncell), dim3(64), 0, st->ctx->stream, A,
                           st->cell_start, st->cell_cnt, st->ctx->g, st->cap, g4);
    else
        hipLaunchKernelGGL((k_diag_deposit<3, T>), dim3((unsigned)st->ncell), dim3(64), 0, st->ctx->stream, A,
                           st->cell_start, st->cell_cnt, st->ctx->g, st->cap, g4);
}

extern "C" int emu_state_flavor_amplitude(emu_state* st, double* sums)
{
    if (!st || !sums) return EMU_ERR_ARG;
    emu_ctx* ctx = st->ctx;
    hipStream_t s = ctx->stream;
    // the auxiliary grid (>= 4 values per cell) is free between steps
    hipLaunchKernelGGL(k_fill_grid, dim3(grid_blocks(st->ncell * st->naux)), dim3(256), 0, s, st->aux,
                       st->ncell * st->naux, 1, 0.0, nullptr, 0);
    if (st->precision == EMU_FP32_BULK) diag_t<float>(st, st->aux);
    else diag_t<double>(st, st->aux);
    // the two sums in red[4..5] (free between steps), as doubles
    hipLaunchKernelGGL(k_fill_u64v, dim3(1), dim3(64), 0, s, st->red + 4, 2, 0ull);
    hipLaunchKernelGGL(k_diag_reduce, dim3(grid_blocks(st->ncell)), dim3(256), 0, s, st->aux, st->ncell,
                       reinterpret_cast<double*>(st->red + 4));
    EMU_LAUNCH_CHECK(ctx, "flavor_amplitude");
    double h[2];
    EMU_HIP(ctx, hipMemcpyAsync(h, st->red + 4, sizeof h, hipMemcpyDeviceToHost, s));
    EMU_HIP(ctx, hipStreamSynchronize(s));
    sums[0] = h[0] / st->nscale;
    sums[1] = h[1] / st->nscale;
    return EMU_OK;
}
