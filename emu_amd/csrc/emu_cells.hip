// emu_cells.hip -- deposit_to_mesh and interpolate_rhs_from_mesh on the
// reference's own data (SoA particles in PIdx order, in any order; the
// `state` MultiFab as one ghost-cell fab), with the particles listed by the
// box their stencil covers.
//
// Each particle's stencil lies in the 3 x 3 x 3 box (one cell wide in the
// dimensions of shape order 0) around its nearest cell: shape3's window.
// A counting sort lists the particles by that box (the key: the fab-box index
// of the box's first cell):
//  - the deposit (Evolve.cpp:91-126, fill generate_code.py:150-169) takes
//    one box per wave and sums D[t][q] = sum_p w_t(p) Q_p,q over the box's
//    particles on the f64 matrix cores (t: the box's cells, q: the 8 NF^2
//    deposited components N f, N f p/p_t for nu and nubar); a workgroup's
//    boxes merge their sums in LDS and add them to the fab with one atomic
//    per value, instead of 27 x 8 NF^2 per particle;
//  - the gather (Evolve.cpp:128-176, fills generate_code.py:239-389) takes
//    two boxes per wave and reads their fields from LDS, where the workgroup
//    stages its run of boxes from a cell-major copy that combines each
//    cell's nu and nubar fields as the gather uses them: per component c,
//    (N + s Nbar + matter delta_c0) and (F_d + s Fbar_d), s = -1 for real and
//    +1 for imaginary parts (generate_code.py:317-323).
//
// The per-particle formulas are the reference's; the order of the sums
// differs (the deposit's also between runs, as with the reference's GPU
// atomics), which moves results by rounding only.  A particle whose box
// leaves the fab is skipped and counted, as emu_sync reports.
#include "emu_cells.h"
#include "emu_stage.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <mutex>
#include <string>

namespace emu {
namespace {

constexpr int kB = 256;

unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kB - 1) / kB, 1 << 22)); }

// Box of a particle in fab-box coordinates: its nearest cell minus one in
// the 3-wide dimensions; false if the box leaves the fab.
__device__ __forceinline__ bool particle_box(const PView& P, int64_t ip, const Fab& fab, const Geom& g, int o[3],
                                             double w[3][3])
{
    bool in = true;
    const double x[3] = {P.pos[0][ip], P.pos[1][ip], P.pos[2][ip]}; // loads issued together
#pragma unroll
    for (int d = 0; d < 3; d++) {
        int idx;
        shape3((x[d] - g.plo[d]) * g.dxi[d], g.order[d], idx, w[d]);
        const int wd = g.order[d] > 0 ? 3 : 1;
        o[d] = idx - (wd == 3 ? 1 : 0) - fab.lo[d];
        in = in && o[d] >= 0 && o[d] + wd <= fab.nf[d];
    }
    return in;
}

// Per-wave aggregation of per-lane keys: for each distinct key among the
// active lanes, f(key, lanes holding it, leader lane).
template <typename F> __device__ __forceinline__ void for_each_key(bool act, int key, F&& f)
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

__global__ void k_box_keys(PView P, Fab fab, Geom g, int nk, int* __restrict__ key, int* __restrict__ counts)
{
    const int64_t ip = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool act = ip < P.np;
    const int lane = threadIdx.x & 63;
    int k = nk;
    if (act) {
        int o[3];
        double w[3][3];
        if (particle_box(P, ip, fab, g, o, w)) k = (o[2] * fab.nf[1] + o[1]) * fab.nf[0] + o[0];
        key[ip] = k;
    }
    for_each_key(act, k, [&](int kk, unsigned long long m, int lead) {
        if (lane == lead) atomicAdd(&counts[kk], __popcll(m));
    });
}

// perm lists the particles box by box; within a box, lane order within a
// wave and waves in the order their atomics land.
__global__ void k_box_perm(int64_t np, const int* __restrict__ key, const int* __restrict__ start, int* cursor,
                           int* __restrict__ perm)
{
    const int64_t ip = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const bool act = ip < np;
    const int lane = threadIdx.x & 63;
    const int k = act ? key[ip] : 0;
    for_each_key(act, k, [&](int kk, unsigned long long m, int lead) {
        int base = 0;
        if (lane == lead) base = atomicAdd(&cursor[kk], __popcll(m));
        base = __shfl(base, lead);
        if ((m >> lane) & 1ull) perm[start[kk] + base + __popcll(m & ((1ull << lane) - 1ull))] = (int)ip;
    });
}

__global__ void k_fill_i32(int* p, int64_t n, int v)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        p[t] = v;
}

__global__ void k_count_skipped(const int* start, int nk, unsigned long long* oob)
{
    const int n = start[nk + 1] - start[nk];
    if (threadIdx.x == 0 && n > 0) atomicAdd(oob, (unsigned long long)n);
}

// The gather's combined fields of every fab cell, cell-major [cell][4][NH]:
// v = 0: N + s Nbar (+ matter on component 0), v = 1..3: F_v + s Fbar_v.
template <int NF> __global__ void k_combine(Fab fab, Geom g, double* __restrict__ comb)
{
    using G = GIdx<NF>;
    constexpr int nh = NH<NF>;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= fab.cstride) return;
    const double* c0 = fab.d + t;
    auto S = [&](int comp) { return c0[(int64_t)comp * fab.cstride]; };
    const double matter = S(G::rho) * S(G::Ye) / phys::Mp / g.inv_cell_volume;
    double* out = comb + t * 4 * nh;
#pragma unroll
    for (int v = 0; v < 4; v++)
#pragma unroll
        for (int c = 0; c < nh; c++) {
            const double a = S(G::var(v, 0) + c), b = S(G::var(v, 1) + c);
            double x = is_im_comp<NF>(c) ? a + b : a - b;
            if (v == 0 && c == 0) x += matter;
            out[v * nh + c] = x;
        }
}

// XCD-aware block order: consecutive boxes share an XCD's L2.
__device__ __forceinline__ int64_t box_block(int blocks_per_xcd)
{
    const int b = blockIdx.x;
    return (int64_t)(b & 7) * blocks_per_xcd + (b >> 3);
}

// The gather takes boxes in runs of kGatherRun along x, one run per
// workgroup and two consecutive boxes per wave.  The run's boxes overlap, so
// the workgroup stages their union once (kGatherRun + 2 cells along x
// instead of 3 per box), and a wave's two boxes fill its lanes better than
// one: 92 particles per box take 3 passes of 64 per pair instead of 4.
constexpr int kGatherRun = 8;
constexpr int kWavesGather = kGatherRun / 2;

template <int NF, int WX, int WY, int WZ>
__global__ __launch_bounds__(64 * kWavesGather) __attribute__((amdgpu_waves_per_eu(5))) void k_gather_boxes(
    PView P, Fab fab, Geom g, const double* __restrict__ comb, const double* __restrict__ M2,
    const int* __restrict__ start, const int* __restrict__ perm, int nrun_x, int64_t nruns, int blocks_per_xcd)
{
    using I = PIdx<NF>;
    constexpr int nh = NH<NF>, QG = 4 * nh, NT = WX * WY * WZ, UX = kGatherRun + WX - 1;
    // the run's fields, the same for every lane of a box: a lane's read of a
    // cell's values is an LDS broadcast (two addresses per wave, one per box)
    __shared__ double2 run_lds[WZ * WY * UX * QG / 2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t run = box_block(blocks_per_xcd);
    if (run >= nruns) return; // the whole workgroup
    const int nf0 = fab.nf[0], nf1 = fab.nf[1];
    const int rx = (int)(run % nrun_x);
    const int64_t ryz = run / nrun_x;
    const int by = (int)(ryz % nf1), bz = (int)(ryz / nf1);
    const int bx0 = rx * kGatherRun;
    const int nbx = min(kGatherRun, nf0 - bx0);                    // boxes in the run
    const int64_t kb0 = ((int64_t)bz * nf1 + by) * nf0 + bx0;      // fab-box index of the first
    if (start[kb0] == start[kb0 + nbx]) return;                    // no particles: the whole workgroup
    {
        // a box with particles lies inside the fab, so rows by..by+WY-1 and
        // bz..bz+WZ-1 exist; along x the union is clamped to the fab
        const int ux = min(nbx + WX - 1, nf0 - bx0);
        const int row2 = ux * QG / 2;
        const int64_t sy = nf0, sz = (int64_t)nf0 * nf1;
        for (int e = threadIdx.x; e < WY * WZ * row2; e += 64 * kWavesGather) {
            const int r = e / row2, wi = e - r * row2;
            const int jj = r % WY, kk = r / WY;
            run_lds[r * (UX * QG / 2) + wi] =
                reinterpret_cast<const double2*>(comb + (kb0 + kk * sz + jj * sy) * QG)[wi];
        }
        __syncthreads();
    }
    if (2 * wave >= nbx) return;
    const int nb = min(2, nbx - 2 * wave);
    const int64_t kb = kb0 + 2 * wave;
    const int beg = __builtin_amdgcn_readfirstlane(start[kb]);
    const int mid = __builtin_amdgcn_readfirstlane(start[kb + 1]);
    const int end = __builtin_amdgcn_readfirstlane(start[kb + nb]);
    if (beg >= end) return;
    const double coef = M_SQRT2 * phys::GF * g.inv_cell_volume;
    const double* runf = reinterpret_cast<const double*>(run_lds);
    for (int base = beg; base < end; base += 64) {
        const bool act = base + lane < end;
        const int64_t ip = perm[act ? base + lane : beg];
        // this lane's box: the wave's first or second
        const double* boxf = runf + (2 * wave + (base + lane >= mid ? 1 : 0)) * QG;
        const double pupx = P.r[I::pupx][ip], pupy = P.r[I::pupy][ip], pupz = P.r[I::pupz][ip],
                     pupt = P.r[I::pupt][ip];
        int o[3];
        double w[3][3];
        particle_box(P, ip, fab, g, o, w);
        const double vx = pupx / pupt, vy = pupy / pupt, vz = pupz / pupt;
        double in[nh];
#pragma unroll
        for (int c = 0; c < nh; c++) in[c] = 0.0;
        // one box cell per iteration: per field v the lane's coefficient
        // w_t s_v, s = (1, -v_x, -v_y, -v_z), times the NH uniform values of
        // that field
        auto pick = [](const double* x, int i) { return i == 0 ? x[0] : (i == 1 ? x[1] : x[2]); };
#pragma unroll 1
        for (int t = 0; t < NT; t++) {
            const int ii = t % WX, jj = (t / WX) % WY, kk = t / (WX * WY);
            const double wt = pick(w[0], WX == 3 ? ii : 1) * pick(w[1], WY == 3 ? jj : 1) * pick(w[2], WZ == 3 ? kk : 1);
            const double* G = boxf + ((kk * WY + jj) * UX + ii) * QG;
            const double a[4] = {wt, -wt * vx, -wt * vy, -wt * vz};
#pragma unroll
            for (int v = 0; v < 4; v++)
#pragma unroll
                for (int c = 0; c < nh; c++) in[c] += a[v] * G[v * nh + c];
        }
        if (!act) continue;
        // vacuum (generate_code.py:239-255), then the gathered potential
        double V[nh], Vb[nh];
#pragma unroll
        for (int i = 0; i < NF; i++)
#pragma unroll
            for (int j = i; j < NF; j++) {
                const int cr = herm_re<NF>(i, j);
                V[cr] = Vb[cr] = M2[cr] * phys::c4 / (2. * pupt);
                if (j > i) {
                    const int ci = herm_im<NF>(i, j);
                    V[ci] = M2[ci] * phys::c4 / (2. * pupt);
                    Vb[ci] = -V[ci];
                }
            }
#pragma unroll
        for (int c = 0; c < nh; c++) {
            const double term = coef * in[c];
            V[c] += term;
            Vb[c] += is_im_comp<NF>(c) ? term : -term;
        }
        // Evolve.cpp:160-172
        P.r[I::x][ip] = pupx / pupt * phys::c;
        P.r[I::y][ip] = pupy / pupt * phys::c;
        P.r[I::z][ip] = pupz / pupt * phys::c;
        P.r[I::time][ip] = 1.0;
        P.r[I::pupx][ip] = 0;
        P.r[I::pupy][ip] = 0;
        P.r[I::pupz][ip] = 0;
        P.r[I::pupt][ip] = 0;
        P.r[I::N][ip] = 0;
        P.r[I::Nbar][ip] = 0;
        P.r[I::L][ip] = 0;
        P.r[I::Lbar][ip] = 0;
        double f[nh], d[nh];
#pragma unroll
        for (int c = 0; c < nh; c++) f[c] = P.r[I::f + c][ip];
        commutator_rhs<NF, double>(V, f, d, 1.0 / phys::hbar);
#pragma unroll
        for (int c = 0; c < nh; c++) P.r[I::f + c][ip] = d[c];
#pragma unroll
        for (int c = 0; c < nh; c++) f[c] = P.r[I::fbar + c][ip];
        commutator_rhs<NF, double>(Vb, f, d, 1.0 / phys::hbar);
#pragma unroll
        for (int c = 0; c < nh; c++) P.r[I::fbar + c][ip] = d[c];
    }
}

// The deposit takes boxes a patch at a time: PX x PY x PZ boxes (2 x 2 x 1 for
// 3-D stencils, 1 x 1 x 4 along z for 1-D ones), one wave each.  Their
// stencil regions overlap, so the waves add their sums into the patch's
// region in LDS first (4 x 4 x 3 cells instead of 4 x 27) and the workgroup
// adds that to the fab: 2.25x fewer device-scope atomics.  Four waves per
// workgroup: the kernel's ~160 VGPRs allow three waves per SIMD, which an
// 8-wave workgroup (two per SIMD) cannot use.
template <int WX, int WY, int WZ> struct DepPatch {
    static constexpr int PX = WX == 3 ? 2 : 1, PY = WY == 3 ? 2 : 1;
    static constexpr int PZ = WZ == 3 ? 4 / (PX * PY) : 1;
    static constexpr int WAVES = PX * PY * PZ;
    static constexpr int RX = PX + WX - 1, RY = PY + WY - 1, RZ = PZ + WZ - 1, NR = RX * RY * RZ;
};

// Particles staged per chunk, and the LDS row of one: the products of its
// shape factors for the box's NT cells (t = ii + WX (jj + WY kk), the matrix
// product's A row), then its sources: N f_c and Nbar fbar_c (2 NH), the
// multipliers (1, p/p_t) of N, Fx, Fy, Fz, and a zero for the padding
// columns.  Both strides are odd, so that four particle rows read in one
// instruction fall on different banks.
constexpr int kDepChunk = 32;
template <int NF, int NT> struct DepRow {
    static constexpr int nh = NH<NF>;
    static constexpr int W = NT;
    static constexpr int S = 2 * nh + 5;
    static constexpr int ONE = 2 * nh, ZERO = 2 * nh + 4;
    static constexpr int doubles_per_wave = kDepChunk * (W + S);
};

template <int NF, int WX, int WY, int WZ>
constexpr int dep_lds_doubles()
{
    using PT = DepPatch<WX, WY, WZ>;
    const int stage = PT::WAVES * DepRow<NF, WX * WY * WZ>::doubles_per_wave, merge = PT::NR * (8 * NH<NF> + 1);
    return stage > merge ? stage : merge;
}

template <int NF, int WX, int WY, int WZ>
__global__ __launch_bounds__((64 * DepPatch<WX, WY, WZ>::WAVES)) __attribute__((amdgpu_waves_per_eu(3))) void k_deposit_boxes(
    PView P, Fab fab, Geom g, const int* __restrict__ start, const int* __restrict__ perm, int npx, int npy,
    int64_t npatch, int blocks_per_xcd)
{
    using I = PIdx<NF>;
    using PT = DepPatch<WX, WY, WZ>;
    constexpr int nh = NH<NF>, NQ = 8 * nh;         // deposited components
    constexpr int NT = WX * WY * WZ;
    using R = DepRow<NF, NT>;
    // full 16-column tiles (f64 16x16x4), then 4-column blocks (f64 4x4x4,
    // four row blocks) for the rest: NF = 3's 72 columns are 4 tiles + 2 blocks
    constexpr int MT = (NT + 15) / 16, NTF = NQ / 16, NSM = (NQ - 16 * NTF + 3) / 4;
    static_assert(PT::NR <= 64, "one lane per region cell in the flush");
    __shared__ double lds[dep_lds_doubles<NF, WX, WY, WZ>()];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t patch = box_block(blocks_per_xcd);
    if (patch >= npatch) return; // the whole workgroup
    const int ppx = (int)(patch % npx), ppy = (int)((patch / npx) % npy), ppz = (int)(patch / ((int64_t)npx * npy));
    // this wave's box, in fab-box coordinates (none past the fab's edge)
    const int wx = wave % PT::PX, wy = (wave / PT::PX) % PT::PY, wz = wave / (PT::PX * PT::PY);
    const int bx = ppx * PT::PX + wx, by = ppy * PT::PY + wy, bz = ppz * PT::PZ + wz;
    const bool have = bx < fab.nf[0] && by < fab.nf[1] && bz < fab.nf[2];
    const int kb = have ? (bz * fab.nf[1] + by) * fab.nf[0] + bx : 0;
    const int beg = have ? __builtin_amdgcn_readfirstlane(start[kb]) : 0;
    const int end = have ? __builtin_amdgcn_readfirstlane(start[kb + 1]) : 0;
    double* Wf = lds + wave * R::doubles_per_wave;
    double* Src = Wf + kDepChunk * R::W;
    // B column q = (2 v + bar) NH + c is source (bar, c) times multiplier v
    // (fab component order, GIdx); columns past NQ read the zero
    auto src1 = [&](int q) { return q < NQ ? ((q / nh) % 2) * nh + q % nh : R::ZERO; };
    auto src2 = [&](int q) { return q < NQ ? 2 * nh + (q / nh) / 2 : R::ZERO; };
    int bi1[NTF], bi2[NTF], bs1[NSM > 0 ? NSM : 1], bs2[NSM > 0 ? NSM : 1];
#pragma unroll
    for (int n = 0; n < NTF; n++) {
        bi1[n] = src1(16 * n + (lane & 15));
        bi2[n] = src2(16 * n + (lane & 15));
    }
#pragma unroll
    for (int k = 0; k < NSM; k++) {
        bs1[k] = src1(16 * NTF + 4 * k + (lane & 3));
        bs2[k] = src2(16 * NTF + 4 * k + (lane & 3));
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
    const int cl = lane & (kDepChunk - 1); // the chunk row this lane stages (lanes 32-63 repeat 0-31)
    const bool row = lane < kDepChunk;
    double* wr = Wf + cl * R::W;
    double* sr = Src + cl * R::S;
    if (row) {
        sr[R::ONE] = 1.0;
        sr[R::ZERO] = 0.0;
    }
    // operand pointers of this lane: particle row 4 ks + lane / 16 of k-step
    // ks; A row t = 16 m + lane % 16 (rows past NT are padding that reads
    // in-range values: row t of the product depends only on row t of A, and
    // those rows are never flushed)
    const double* wl = Wf + (lane >> 4) * R::W + (lane & 15);
    const double* sl = Src + (lane >> 4) * R::S;
    // each chunk's permutation entry is loaded one chunk ahead, so that the
    // particle loads wait on one round trip instead of two
    int ip_next = beg < end ? perm[beg + cl < end ? beg + cl : beg] : 0;
    for (int base = beg; base < end; base += kDepChunk) {
        const bool act = base + cl < end;
        const int64_t ip = ip_next;
        const int nb = base + kDepChunk;
        if (nb < end) ip_next = perm[nb + cl < end ? nb + cl : beg];
        int o[3];
        double w[3][3];
        particle_box(P, ip, fab, g, o, w);
        // every lane loads (inactive ones a valid particle), so that the loads
        // issue together, nu's batch then nubar's.  An inactive row gets zero
        // weights: its sources then add nothing.
        {
            const double pupt = P.r[I::pupt][ip], pv[3] = {P.r[I::pupx][ip], P.r[I::pupy][ip], P.r[I::pupz][ip]};
            const double N = P.r[I::N][ip];
            double fv[nh];
#pragma unroll
            for (int c = 0; c < nh; c++) fv[c] = P.r[I::f + c][ip];
            if (row) {
                auto pick = [](const double* x, int i) { return i == 0 ? x[0] : (i == 1 ? x[1] : x[2]); };
#pragma unroll
                for (int kk = 0; kk < WZ; kk++)
#pragma unroll
                    for (int jj = 0; jj < WY; jj++) {
                        const double yz = pick(w[1], WY == 3 ? jj : 1) * pick(w[2], WZ == 3 ? kk : 1);
#pragma unroll
                        for (int ii = 0; ii < WX; ii++) {
                            const double x = pick(w[0], WX == 3 ? ii : 1);
                            wr[(kk * WY + jj) * WX + ii] = (act ? x : 0.0) * yz;
                        }
                    }
#pragma unroll
                for (int c = 0; c < nh; c++) sr[c] = N * fv[c];
                const double ipt = 1.0 / pupt;
#pragma unroll
                for (int d = 0; d < 3; d++) sr[R::ONE + 1 + d] = pv[d] * ipt;
            }
        }
        {
            const double Nb = P.r[I::Nbar][ip];
            double fv[nh];
#pragma unroll
            for (int c = 0; c < nh; c++) fv[c] = P.r[I::fbar + c][ip];
            if (row) {
#pragma unroll
                for (int c = 0; c < nh; c++) sr[nh + c] = Nb * fv[c];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // k-steps of 4 particles, two per iteration (the second at immediate
        // offsets); an odd count runs one more k-step over zero-weight rows
        int nks = (min(kDepChunk, end - base) + 3) >> 2;
        nks += nks & 1;
        for (int ks = 0; ks < nks; ks += 2) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const double* wk = wl + (ks + h) * 4 * R::W;
                const double* sk = sl + (ks + h) * 4 * R::S;
                double A[MT];
#pragma unroll
                for (int m = 0; m < MT; m++) A[m] = wk[16 * m];
#pragma unroll
                for (int n = 0; n < NTF; n++) {
                    const double Bv = sk[bi1[n]] * sk[bi2[n]];
#pragma unroll
                    for (int m = 0; m < MT; m++)
                        acc[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[m], Bv, acc[m][n], 0, 0, 0);
                }
#pragma unroll
                for (int k = 0; k < NSM; k++) {
                    const double Bs = sk[bs1[k]] * sk[bs2[k]];
#pragma unroll
                    for (int m = 0; m < MT; m++)
                        accs[m][k] = __builtin_amdgcn_mfma_f64_4x4x4f64(A[m], Bs, accs[m][k], 0, 0, 0);
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // ---- merge the patch's boxes in LDS (the staging rows are free now), then
    // add the region to the fab; the region's first cell is the patch's first box
    constexpr int NQS = NQ + 1; // odd row stride: a wave reading one component over the cells spreads over the banks
    double* Mb = lds; // [NR][NQS]
    __syncthreads();
    for (int e = threadIdx.x; e < PT::NR * NQS; e += 64 * PT::WAVES) Mb[e] = 0.0;
    __syncthreads();
    if (beg < end) {
        // element r of lane l holds D[16m + 4r + l/16][16n + l%16]
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int t = 16 * m + 4 * r + (lane >> 4);
                if (t >= NT) continue;
                const int ii = t % WX, jj = (t / WX) % WY, kk = t / (WX * WY);
                const int rc = ((wz + kk) * PT::RY + (wy + jj)) * PT::RX + (wx + ii);
#pragma unroll
                for (int n = 0; n < NTF; n++) unsafeAtomicAdd(&Mb[rc * NQS + 16 * n + (lane & 15)], acc[m][n][r]);
            }
        // 4x4x4 blocks: element of lane l is D[16m + 4((l >> 2) & 3) + l/16][16 NTF + 4k + l%4]
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int k = 0; k < NSM; k++) {
                const int t = 16 * m + 4 * ((lane >> 2) & 3) + (lane >> 4), q = 16 * NTF + 4 * k + (lane & 3);
                if (t >= NT || q >= NQ) continue;
                const int ii = t % WX, jj = (t / WX) % WY, kk = t / (WX * WY);
                const int rc = ((wz + kk) * PT::RY + (wy + jj)) * PT::RX + (wx + ii);
                unsafeAtomicAdd(&Mb[rc * NQS + q], accs[m][k]);
            }
    }
    __syncthreads();
    // lane l takes region cell l, and each wave a quarter of the components:
    // a wave's atomics per pass fall on the region's cells of one fab component
    {
        const int rc = lane;
        const int cx = ppx * PT::PX + rc % PT::RX, cy = ppy * PT::PY + (rc / PT::RX) % PT::RY;
        const int cz = ppz * PT::PZ + rc / (PT::RX * PT::RY);
        if (rc < PT::NR && cx < fab.nf[0] && cy < fab.nf[1] && cz < fab.nf[2]) {
            double* dst = fab.d + 3 * fab.cstride + ((int64_t)cz * fab.nf[1] + cy) * fab.nf[0] + cx;
            const double* srcm = Mb + rc * NQS;
            for (int q = wave; q < NQ; q += PT::WAVES) {
                const double v = srcm[q];
                if (v != 0.0) unsafeAtomicAdd(dst + (int64_t)q * fab.cstride, v);
            }
        }
    }
}

template <int NF, int WX, int WY, int WZ>
void launch_boxes(bool deposit, hipStream_t s, const PView& P, const Fab& fab, const Geom& g, const double* comb,
                  const double* M2, const int* start, const int* perm, int nk)
{
    if (deposit) {
        using PT = DepPatch<WX, WY, WZ>;
        const int npx = (fab.nf[0] + PT::PX - 1) / PT::PX, npy = (fab.nf[1] + PT::PY - 1) / PT::PY;
        const int npz = (fab.nf[2] + PT::PZ - 1) / PT::PZ;
        const int64_t npatch = (int64_t)npx * npy * npz;
        const int bpx = (int)((npatch + 7) / 8);
        hipLaunchKernelGGL((k_deposit_boxes<NF, WX, WY, WZ>), dim3(8 * bpx), dim3(64 * PT::WAVES), 0, s, P, fab, g,
                           start, perm, npx, npy, npatch, bpx);
    } else {
        const int nrun_x = (fab.nf[0] + kGatherRun - 1) / kGatherRun;
        const int64_t nruns = (int64_t)nrun_x * fab.nf[1] * fab.nf[2];
        const int bpx = (int)((nruns + 7) / 8);
        hipLaunchKernelGGL((k_gather_boxes<NF, WX, WY, WZ>), dim3(8 * bpx), dim3(64 * kWavesGather), 0, s, P, fab, g,
                           comb, M2, start, perm, nrun_x, nruns, bpx);
    }
}

template <int NF>
void launch_boxes_nf(bool deposit, const Geom& g, hipStream_t s, const PView& P, const Fab& fab, const double* comb,
                     const double* M2, const int* start, const int* perm, int nk)
{
    const int w = (g.order[0] > 0 ? 1 : 0) | (g.order[1] > 0 ? 2 : 0) | (g.order[2] > 0 ? 4 : 0);
    switch (w) {
    case 0: launch_boxes<NF, 1, 1, 1>(deposit, s, P, fab, g, comb, M2, start, perm, nk); break;
    case 1: launch_boxes<NF, 3, 1, 1>(deposit, s, P, fab, g, comb, M2, start, perm, nk); break;
    case 2: launch_boxes<NF, 1, 3, 1>(deposit, s, P, fab, g, comb, M2, start, perm, nk); break;
    case 3: launch_boxes<NF, 3, 3, 1>(deposit, s, P, fab, g, comb, M2, start, perm, nk); break;
    case 4: launch_boxes<NF, 1, 1, 3>(deposit, s, P, fab, g, comb, M2, start, perm, nk); break;
    case 5: launch_boxes<NF, 3, 1, 3>(deposit, s, P, fab, g, comb, M2, start, perm, nk); break;
    case 6: launch_boxes<NF, 1, 3, 3>(deposit, s, P, fab, g, comb, M2, start, perm, nk); break;
    default: launch_boxes<NF, 3, 3, 3>(deposit, s, P, fab, g, comb, M2, start, perm, nk); break;
    }
}

size_t scan_bytes(int64_t n)
{
    size_t b = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, b, (const int*)nullptr, (int*)nullptr, (int)n);
    return b;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

} // namespace

BoxScratchSize box_scratch_size(int64_t np, int64_t nk, int nf, bool gather)
{
    BoxScratchSize s;
    s.ints = align256(sizeof(int) * (size_t)(2 * np + 3 * (nk + 2)));
    s.comb = gather ? align256(sizeof(double) * (size_t)nk * 4 * nf * nf) : 0;
    s.scan = align256(scan_bytes(nk + 2));
    s.total = s.ints + s.comb + s.scan;
    return s;
}

int box_list_and_run(emu_ctx* ctx, const PView& P, const Fab& fab, bool deposit, void* scratch, size_t scratch_bytes)
{
    hipStream_t s = ctx->stream;
    const int64_t nk64 = fab.cstride;
    if (nk64 + 2 > (int64_t)INT32_MAX || P.np > (int64_t)INT32_MAX)
        return set_error(ctx, EMU_ERR_UNSUPPORTED, "more than 2^31 particles or fab cells");
    const int nk = (int)nk64;
    const int nf = ctx->p.nf;
    const BoxScratchSize sz = box_scratch_size(P.np, nk, nf, !deposit);
    if (scratch_bytes < sz.total) return set_error(ctx, EMU_ERR_ARG, "box list scratch too small");
    char* b = (char*)scratch;
    int* key = (int*)b;
    int* perm = key + P.np;
    int* counts = perm + P.np;
    int* start = counts + (nk + 2);
    int* cursor = start + (nk + 2);
    double* comb = (double*)(b + sz.ints);
    void* tmp = b + sz.ints + sz.comb;
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(2 * (nk + 2))), dim3(kB), 0, s, counts, 2 * (int64_t)(nk + 2), 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(nk + 2)), dim3(kB), 0, s, cursor, (int64_t)(nk + 2), 0);
    if (P.np > 0)
        hipLaunchKernelGGL(k_box_keys, dim3(nblk(P.np)), dim3(kB), 0, s, P, fab, ctx->g, nk, key, counts);
    size_t tb = sz.scan;
    EMU_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, tb, counts, start, nk + 2, s));
    if (P.np > 0) hipLaunchKernelGGL(k_box_perm, dim3(nblk(P.np)), dim3(kB), 0, s, P.np, key, start, cursor, perm);
    hipLaunchKernelGGL(k_count_skipped, dim3(1), dim3(64), 0, s, start, nk, ctx->d_red + 7);
    if (!deposit) {
        if (nf == 2) hipLaunchKernelGGL(k_combine<2>, dim3(nblk(nk)), dim3(kB), 0, s, fab, ctx->g, comb);
        else hipLaunchKernelGGL(k_combine<3>, dim3(nblk(nk)), dim3(kB), 0, s, fab, ctx->g, comb);
    }
    if (P.np > 0) {
        if (nf == 2) launch_boxes_nf<2>(deposit, ctx->g, s, P, fab, comb, ctx->d_M2, start, perm, nk);
        else launch_boxes_nf<3>(deposit, ctx->g, s, P, fab, comb, ctx->d_M2, start, perm, nk);
    }
    EMU_LAUNCH_CHECK(ctx, deposit ? "deposit_to_mesh" : "interpolate_rhs_from_mesh");
    return EMU_OK;
}

// ---------------------------------------------------------------------------
// Scratch for the deposit that allocates nothing (emu_deposit_to_mesh_static):
// one array per GPU, in the code object's device memory, for at most
// kStaticMaxParticles particles and kStaticMaxCells fab cells (ghosts
// included).  Host code uses the per-device address from hipGetSymbolAddress,
// never the array's name.  Every call on one GPU shares the array.
// ---------------------------------------------------------------------------
constexpr int64_t kStaticMaxParticles = int64_t(1) << 22;
constexpr int64_t kStaticMaxCells = int64_t(1) << 20;
constexpr size_t kStaticScanBytes = size_t(1) << 20;
constexpr size_t kStaticScratchBytes =
    ((sizeof(int) * (size_t)(2 * kStaticMaxParticles + 3 * (kStaticMaxCells + 2)) + 255) & ~(size_t)255) +
    kStaticScanBytes;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
static void* g_static_scratch[kMaxDevices] = {};
static std::mutex g_static_mu;

int static_scratch_init()
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return EMU_ERR_HIP;
    std::lock_guard<std::mutex> lock(g_static_mu);
    if (g_static_scratch[dev]) return EMU_OK;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return EMU_ERR_HIP;
    g_static_scratch[dev] = p;
    return EMU_OK;
}

int static_scratch(emu_ctx* ctx, int64_t np, int64_t ncells, void** out, size_t* bytes)
{
    if (np > kStaticMaxParticles || ncells > kStaticMaxCells)
        return set_error(ctx, EMU_ERR_UNSUPPORTED,
                         "emu_deposit_to_mesh_static holds at most " + std::to_string(kStaticMaxParticles) +
                             " particles and " + std::to_string(kStaticMaxCells) + " fab cells (got " +
                             std::to_string(np) + " and " + std::to_string(ncells) + ")");
    const BoxScratchSize sz = box_scratch_size(np, ncells, ctx->p.nf, false);
    if (sz.total > kStaticScratchBytes || sz.scan > kStaticScanBytes)
        return set_error(ctx, EMU_ERR_UNSUPPORTED, "static scratch too small for this grid");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices || !g_static_scratch[dev])
        return set_error(ctx, EMU_ERR_HIP, "static scratch not set up on this device (emu_create sets it up)");
    *out = g_static_scratch[dev];
    *bytes = sz.total;
    return EMU_OK;
}

int box_scratch(emu_ctx* ctx, size_t need, void** out)
{
    if (ctx->box_scratch_bytes < need) {
        if (ctx->box_scratch) (void)hipFree(ctx->box_scratch);
        ctx->box_scratch = nullptr;
        ctx->box_scratch_bytes = 0;
        if (hipMalloc(&ctx->box_scratch, need) != hipSuccess) {
            ctx->box_scratch = nullptr;
            return set_error(ctx, EMU_ERR_NOMEM, "box list scratch");
        }
        ctx->box_scratch_bytes = need;
    }
    *out = ctx->box_scratch;
    return EMU_OK;
}

} // namespace emu
