// emu_stage.h -- the fused stage kernel: gather + Hamiltonian + commutator +
// RK update of every particle of a cell, and the deposit of the next stage
// state as one small matrix product per cell on the FP64 matrix cores.
//
// One wave owns one cell (FP32-bulk's RK stages: two x-adjacent cells) and
// takes its particles 64 at a time (lane = particle).
//
// Gather (MFMA): with s(p) = (1, -vx, -vy, -vz) and G_g(t, c) the four field
// groups (M, Phix, Phiy, Phiz) of neighbour t,
//   inside_c(p) = sum_t sum_g [w_t(p) s_g(p)] G_g(t, c),
// a (16 particles x 4*27) x (4*27 x NH) product per tile of 16 particles: one
// 16x16x4 step per neighbour t, its K = 4 slice being the four groups.  The
// cell's 27 x 4 NH neighbour fields are staged in LDS once per cell with
// coalesced loads and feed the B operand; A is the particle's weight times
// s_g, read from the shape factors staged in LDS.
// Deposit (MFMA): the contribution of the cell's particles p to target cell t
// and field column q is D[t][q] = sum_p W[p][t] R[p][q], with W the shape
// weights and R = (1, vx, vy, vz) x q_c(p), q = N f - conj(Nbar fbar).  Over a
// chunk of 64 particles this is a (27 x 64) x (64 x 4NH) product, done with
// v_mfma_f64_16x16x4 (4 particles per instruction), so the sum over particles
// never leaves the matrix unit; the lane-layout transposition goes through LDS
// (9 shape factors and NH + 3 sources per particle).  The wave's accumulators
// are added to the grid once per cell with FP64 atomics.
#pragma once

#include "emu_internal.h"

namespace emu {

typedef double mf_d4 __attribute__((ext_vector_type(4)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));

// One 16x16x4 matrix-core step in the precision of T.  Element r of lane l of
// the result holds row 4r + l/16 (f64) or 4(l/16) + r (f32), column l%16.
template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef mf_d4 acc_t;
    static __device__ __forceinline__ acc_t step(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int r, int lane) { return 4 * r + (lane >> 4); }
};
template <> struct Mfma<float> {
    typedef mf_f4 acc_t;
    static __device__ __forceinline__ acc_t step(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int r, int lane) { return 4 * (lane >> 4) + r; }
};

template <typename T> struct PArrays {
    double* pos[3];
    int64_t* id;
    T* pup[4]; // pupx, pupy, pupz, pupt
    T* Nn[2];  // N, Nbar
    T* Ll[2];  // L, Lbar
    T* f;      // [2 NH][np]: nu components then nubar
};

template <typename T> struct StageArgs {
    PArrays<T> cur; // current copy, sorted by cell_start
    PArrays<T> dst; // stage 4 writes the step's result here, sorted by the new cells
    T* Y;           // stage state
    T* acc;         // RK accumulator
    int* newcell;   // stage 1 writes each particle's end-of-step cell
    const int* mbase;
    const int* rank;
    const unsigned char* mv;
    const int* cell_start; // first slot of each cell (a multiple of the slot alignment)
    const int* cell_cnt;   // particles of each cell
    const double* gin;  // grid read by the gather
    double* gout;       // grid deposited by this launch
    double* aux;        // compute_dt auxiliaries
    const double* dt;
    unsigned long long* err;
    int64_t np; // slots per array: the stride of the component-major arrays
    int64_t ncell;
    int stage; // 0 = deposit only, 1..4
    int want_aux;
    int blocks_per_xcd;
    Geom g;
    double vac[9]; // M2 c^4 / (2 hbar)
    double coef;   // sqrt2 GF / dV / hbar
    double maxError;
};

template <int NF> __device__ __forceinline__ constexpr bool is_im_comp(int c)
{
    for (int a = 0; a < NF; a++)
        for (int b = a + 1; b < NF; b++)
            if (c == herm_im<NF>(a, b)) return true;
    return false;
}

__device__ __forceinline__ int bin1(double pos, double plo, double dxi, int n)
{
#pragma clang fp contract(off)
    return pmod(nearest_cell_center_index((pos - plo) * dxi), n);
}

// x += (dt b_i) v for the four RK4 stages, rounded as the reference's saxpy
// sequence (FlavoredNeutrinoContainer.H:43-45) with v = c pup/pupt.
__device__ __forceinline__ double advance_pos(double x, double v, const double a[4])
{
#pragma clang fp contract(off)
    x = a[0] * v + x;
    x = a[1] * v + x;
    x = a[2] * v + x;
    x = a[3] * v + x;
    return x;
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

template <int NF>
__device__ __forceinline__ int renorm_fp64(double* f, double L, double maxError)
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

// End-of-step position: the four stage saxpys, then the periodic shift.
__device__ __forceinline__ void end_position(const double pos[3], const double v[3], double dt, const Geom& g,
                                             double npos[3], int& newcell)
{
    const double ai[4] = {dt * (1. / 6.), dt * (1. / 3.), dt * (1. / 3.), dt * (1. / 6.)};
    int nc[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        npos[d] = wrap1(advance_pos(pos[d], v[d], ai), g.rlo[d], g.rhi[d], g.L[d]);
        nc[d] = bin1(npos[d], g.plo[d], g.dxi[d], g.n[d]);
    }
    newcell = (nc[2] * g.n[1] + nc[1]) * g.n[0] + nc[0];
}

// Three ways to map waves to cells (MODE).
//  0: every wave is its own patch of one cell (box = its 27 neighbours; two
//     cells in FP32-bulk, below), stages its own box and flushes its deposit
//     directly: no workgroup barrier.
//  1: the WPB waves of a workgroup take a patch of WPB cells (2 along x, then
//     2 along y, when the stencil is 3 wide there, else along z); the patch's
//     neighbourhood box (PX + WX - 1) x (PY + WY - 1) x (PZ + WZ - 1) is staged
//     once for all of them (one barrier), and each wave flushes directly.
//  2: as 1, and the deposits are also summed over the box in LDS before one
//     set of global atomics (56% fewer for a 2 x 2 patch at the 3-D stencil),
//     at the cost of two more barriers.
// PM (a shared patch) is MODE >= 1.  In MODE 0 a wave may also take CPW = 2
// cells, x-adjacent, as one patch: their particles form one list of 16-particle
// tiles (the first cell's, then the second's), so that a cell of 92 particles no
// longer leaves 36 of a chunk's 64 lanes idle, and the wave stages the two
// cells' 4 x 3 x 3 box once.
template <int WX, int WY, int WZ, int WPB, bool PM, int CPW = 1> struct Patch {
    static_assert(CPW == 1 || (CPW == 2 && !PM && WX == 3), "two cells per wave: MODE 0, along a 3-wide x");
    static constexpr int PX = (PM || CPW == 2) && WX == 3 ? 2 : 1;
    static constexpr int PY = PM && WY == 3 && WPB >= 2 * PX ? 2 : 1;
    static constexpr int PZ = PM ? WPB / (PX * PY) : 1;
    static constexpr int CELLS = PM ? WPB : CPW; // cells of one patch
    static_assert(PX * PY * PZ == CELLS, "a patch holds one cell per wave, or the wave's CPW cells");
    static constexpr int NBX = PX + WX - 1, NBY = PY + WY - 1, NBZ = PZ + WZ - 1;
    static constexpr int NB = NBX * NBY * NBZ;
};

// compute_dt auxiliaries per cell.  AUX = 1: the diagonal sums N_ii + Nbar_ii
// (i < NF), all that the conservative flavour bound needs; AUX = 2: also
// Nbar Im(fbar_ab) v_d per pair a < b and direction d, for the adaptive bound
// (max_adaptive_speedup > 1).  Deposited columns, and sources per particle:
// N f_ii + Nbar fbar_ii, then Nbar Im(fbar_ab).
template <int NF> constexpr int aux_columns(int AUX)
{
    return AUX == 0 ? 0 : (AUX == 1 ? NF : NF + 3 * (NF * (NF - 1) / 2));
}
template <int NF> constexpr int aux_sources(int AUX)
{
    return AUX == 0 ? 0 : (AUX == 1 ? NF : NF + NF * (NF - 1) / 2);
}

// Values of T per particle in the LDS source rows: NH + 4 main sources, the
// auxiliary ones, and one of padding where that count is even.  The deposit
// reads one column of four consecutive rows at once; an odd row length keeps
// those rows off the same LDS banks (16 doubles = 128 bytes would put every
// row on bank 0).
template <int NF> constexpr int stage_source_row(int AUX)
{
    const int n = NF * NF + 4 + aux_sources<NF>(AUX);
    return n % 2 == 0 ? n + 1 : n;
}

// LDS of one wave's own region, in doubles: 9 shape factors per particle
// (double) and the source row (T).
template <int NF, typename T, int AUX> constexpr int stage_lds_doubles_per_wave()
{
    return 64 * 9 + 64 * stage_source_row<NF>(AUX) * (int)sizeof(T) / 8;
}

// LDS of the table of the neighbourhood box's cell indices (NB ints), in
// doubles: one per workgroup (PM) or per wave.
template <int WX, int WY, int WZ, int WPB, bool PM, int CPW> constexpr int stage_lds_ctab_doubles()
{
    return (Patch<WX, WY, WZ, WPB, PM, CPW>::NB + 1) / 2;
}

// cells per wave: two in FP32-bulk's independent-wave launches of the 3-D
// stencil (its 4 x 3 x 3 box still fits three waves per SIMD in LDS; FP64's
// would not fit two)
template <typename T, int WX, int MODE> constexpr int stage_cells_per_wave()
{
    return MODE == 0 && sizeof(T) == 4 && WX == 3 ? 2 : 1;
}

// LDS of a workgroup, in doubles.  The cell tables come first.  PM: then the
// staged neighbourhood box (gather stages, NB x 4 NH values of T) and the
// waves' regions; with the merge (MODE 2) the deposit's merge buffer (NB x
// deposited columns, double) reuses that space once every wave is done with
// it.  Without PM: per wave, its own box and its region.  The host sizes the
// launch with this function.
template <int NF, typename T, int WX, int WY, int WZ, int WPB, int S, int AUX, int MODE>
constexpr int stage_lds_doubles_per_block()
{
    constexpr bool PM = MODE >= 1;
    constexpr int CPW = stage_cells_per_wave<T, WX, MODE>();
    constexpr int nh = NF * NF;
    constexpr int NB = Patch<WX, WY, WZ, WPB, PM, CPW>::NB;
    constexpr int ct = (PM ? 1 : WPB) * stage_lds_ctab_doubles<WX, WY, WZ, WPB, PM, CPW>();
    constexpr int gsh = S == 0 ? 0 : (NB * 4 * nh * (int)sizeof(T) + 7) / 8;
    if (!PM) return ct + WPB * (gsh + stage_lds_doubles_per_wave<NF, T, AUX>());
    constexpr int work = gsh + WPB * stage_lds_doubles_per_wave<NF, T, AUX>();
    constexpr int merge = MODE == 2 && S <= 3 ? NB * (4 * nh + aux_columns<NF>(AUX)) : 0;
    return ct + (work > merge ? work : merge);
}

// the mapping of each launch kind (measured per launch kind at 64^3 x 92,
// NF = 3): the deposit-only launch in both precisions and FP32-bulk stage 4
// merge their deposits (MODE 2); the other RK stages run faster with
// independent waves
template <typename T, int S, int AUX> constexpr int stage_mode()
{
    return S == 0 || (S == 4 && sizeof(T) == 4) ? 2 : 0;
}
// waves per workgroup: a patch of 4 (MODE 2); otherwise 2, or 1 for the launches
// with the compute_dt auxiliaries, whose larger LDS rows would otherwise leave
// a CU a whole 2-wave workgroup short (FP64 at 64^3 x 92: 6 waves per CU with
// 2-wave workgroups, 7 with single waves)
template <int MODE, int AUX> constexpr int stage_waves_per_block() { return MODE == 2 ? 4 : (AUX != 0 ? 1 : 2); }

// S: 0 = deposit only (the step's result, or the initial state), 1..4 = RK
// stage; AUX: also deposit the compute_dt auxiliaries (S = 0 or 3; 1 or 2 as
// for aux_columns).
template <int NF, typename T, int WX, int WY, int WZ, int WPB, int S, int AUX, int MODE>
__global__ __launch_bounds__(64 * WPB) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 3 : 2))) void k_stage(StageArgs<T> a)
{
    constexpr bool PM = MODE >= 1;    // the workgroup shares one patch and its staged box
    constexpr bool MERGE = MODE == 2; // and merges its deposits in LDS
    constexpr int nh = NH<NF>;
    constexpr int QG = 4 * nh;
    constexpr int NT = WX * WY * WZ;        // target cells of a source cell
    constexpr int MT = (NT + 15) / 16;       // MFMA row tiles
    constexpr int QGA = QG + aux_columns<NF>(AUX); // deposited columns
    constexpr int NTL = (QGA + 15) / 16;     // MFMA column tiles
    constexpr int NPAIR = NF * (NF - 1) / 2;
    // compute_dt auxiliaries per cell (aux_columns): N_ii + Nbar_ii, then (AUX = 2)
    // Nbar Im(fbar_ab) v_d.  N_ii and Nbar_ii separately follow from the sum and
    // M_ii = N_ii - Nbar_ii (+ matter on 00).  They are deposited as columns
    // QG .. QGA - 1 of the same product, into rows of NAUX (the full set's
    // count, whichever part is deposited).
    constexpr int NAUX = NF + 3 * NPAIR;
    constexpr int NSRC = nh + 4;             // q_c, vhat and the constant 1 per particle
    constexpr int NSAUX = aux_sources<NF>(AUX);
    constexpr int NSR = stage_source_row<NF>(AUX); // a particle's source row: main, auxiliary, padding
    static_assert(NSR >= NSRC + NSAUX, "source row");
    extern __shared__ double lds[];

    const Geom& g = a.g;
    constexpr int s = S;
    constexpr bool aux = AUX != 0;
    constexpr int CPW = stage_cells_per_wave<T, WX, MODE>();
    using PT = Patch<WX, WY, WZ, WPB, PM, CPW>;
    constexpr int NB = PT::NB;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    // blocks go round-robin over the 8 XCDs; give each XCD a contiguous run of patches
    const int npx = (g.n[0] + PT::PX - 1) / PT::PX, npy = (g.n[1] + PT::PY - 1) / PT::PY;
    const int npz = (g.n[2] + PT::PZ - 1) / PT::PZ;
    const int b = blockIdx.x;
    const int bgroup = (b & 7) * a.blocks_per_xcd + (b >> 3);
    // PM: the block's patch; otherwise this wave's patch of CPW cells, WPB consecutive per block
    const int patch = __builtin_amdgcn_readfirstlane(PM ? bgroup : bgroup * WPB + wave);
    if (patch >= npx * npy * npz) return; // uniform over the workgroup (PM) or the wave
    const int bx = (patch % npx) * PT::PX, by = ((patch / npx) % npy) * PT::PY, bz = (patch / (npx * npy)) * PT::PZ;
    const int pw = PM ? wave : 0; // this wave's cell within the patch
    const int ox = pw % PT::PX, oy = (pw / PT::PX) % PT::PY, oz = pw / (PT::PX * PT::PY);
    const int cx = bx + ox, cy = by + oy, cz = bz + oz;
    // a wave whose cell lies outside the domain (a partial patch) or holds no
    // particle has nothing to do, but it stays for the workgroup's barriers
    const bool inside_domain = cx < g.n[0] && cy < g.n[1] && cz < g.n[2];
    const int cell = inside_domain ? (cz * g.n[1] + cy) * g.n[0] + cx : 0;
    // (wave-uniform, so the particle arrays are addressed from scalar bases)
    const int beg = __builtin_amdgcn_readfirstlane(inside_domain ? a.cell_start[cell] : 0);
    const int end = __builtin_amdgcn_readfirstlane(inside_domain ? beg + a.cell_cnt[cell] : 0);
    // CPW = 2: the wave's second cell, x + 1 (none past the domain's last column)
    const bool inside1 = CPW == 2 && inside_domain && cx + 1 < g.n[0];
    const int cell1 = inside1 ? cell + 1 : 0;
    const int beg1 = __builtin_amdgcn_readfirstlane(inside1 ? a.cell_start[cell1] : 0);
    const int end1 = __builtin_amdgcn_readfirstlane(inside1 ? beg1 + a.cell_cnt[cell1] : 0);
    // cell i of the neighbourhood box, box index (i, j, k) <-> the patch base
    // minus one in each 3-wide dimension, plus (i, j, k), periodic
    // (one conditional wrap covers a single cell's box: offsets are -1..+1 and a
    // 3-wide stencil implies at least 2 cells in that dimension)
    auto wrapc = [](int x, int n) {
        return PM || CPW == 2 ? pmod(x, n) : (x < 0 ? x + n : (x >= n ? x - n : x));
    };
    auto box_cell = [&](int e) -> int64_t {
        const int i = e % PT::NBX, j = (e / PT::NBX) % PT::NBY, k = e / (PT::NBX * PT::NBY);
        const int ci = wrapc(bx - (WX == 3 ? 1 : 0) + i, g.n[0]);
        const int cj = wrapc(by - (WY == 3 ? 1 : 0) + j, g.n[1]);
        const int ck = wrapc(bz - (WZ == 3 ? 1 : 0) + k, g.n[2]);
        return ((int64_t)ck * g.n[1] + cj) * g.n[0] + ci;
    };
    // box index of this wave's stencil target t (of its first cell; the second
    // cell's is one further along x)
    auto box_of = [&](int t) -> int {
        const int ii = t % WX, jj = (t / WX) % WY, kk = t / (WX * WY);
        return ((oz + kk) * PT::NBY + (oy + jj)) * PT::NBX + (ox + ii);
    };
    constexpr int GSH = s == 0 ? 0 : (NB * QG * (int)sizeof(T) + 7) / 8; // doubles of the staged box
    constexpr int per_wave = stage_lds_doubles_per_wave<NF, T, AUX>();
    static_assert(per_wave == 64 * 9 + 64 * NSR * (int)sizeof(T) / 8,
                  "LDS layout and its size function disagree");
    constexpr int CT = stage_lds_ctab_doubles<WX, WY, WZ, WPB, PM, CPW>();
    // [NB] the box's cell indices: the block's, or this wave's
    int* ctab = (int*)(PM ? lds : lds + wave * CT);
    double* lw = lds + (PM ? CT : WPB * CT);
    // [NB][QG] neighbourhood fields (gather stages): the block's, or this wave's
    T* Gs = (T*)(PM ? lw : lw + wave * (GSH + per_wave));
    double* Wf = PM ? lw + GSH + wave * per_wave : lw + wave * (GSH + per_wave) + GSH; // [64][9] shape factors
    // q_c, vhat (inside_c during the gather): values of the working precision T,
    // so FP32-bulk keeps them as float (lossless) and the LDS per wave shrinks
    T* Src = (T*)(Wf + 64 * 9);             // [64][NSR]
    // the auxiliary sources sit in T in the same rows: in FP32-bulk the products
    // N f are FP32 values already (their sum over particles is accumulated in FP64)
    constexpr int STRIDE = PM ? 64 * WPB : 64;
    const int tid = PM ? (int)threadIdx.x : lane;
    // the box's cell indices, once: the staging loads and the deposit's flush
    // read them instead of redoing the periodic index arithmetic per element
    for (int e = tid; e < NB; e += STRIDE) ctab[e] = (int)box_cell(e);
    if constexpr (PM) {
        __syncthreads();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if constexpr (s != 0) {
        // stage the box's fields, all waves together: NB rows of QG values, in
        // the working precision (FP32-bulk rounds the FP64 deposit here)
        const double* gsrc = a.gin;
        // a compile-time trip count, unrolled, so that every load is in flight at once
#pragma unroll
        for (int e0 = 0; e0 < NB * QG; e0 += STRIDE) {
            const int e = e0 + tid;
            if (e < NB * QG) {
                const int r = e / QG, q = e - r * QG;
                Gs[e] = T(gsrc[(int64_t)ctab[r] * QG + q]);
            }
        }
        if constexpr (PM) {
            __syncthreads();
        } else {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }

    const double dt = *a.dt;
    const double bw[4] = {1. / 6., 1. / 3., 1. / 3., 1. / 6.};
    const double cnext[4] = {0.5, 0.5, 1.0, 0.0};

    // lane constants of the deposit's MFMA operands.  A (row t = 16m + l%16):
    // the three shape factors of target t.  B (column q = 16n + l%16): the
    // product of two entries of the particle's source row, the constant 1
    // standing in where a column has one factor.  Row t of the product depends
    // only on row t of A and column q only on column q of B, so the padding
    // rows (t >= NT) and columns (q >= QGA), which are never flushed, read any
    // in-range entries instead of being masked.
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
    // The last columns, when at most 8 remain past the 16-column tiles (NF = 3:
    // 36 = 2 x 16 + 4 without the auxiliaries, 39 = 2 x 16 + 7 with the
    // diagonal ones), run as 4-column tiles on the 4-block 4x4x4 form instead:
    // 16-cycle steps in place of a 64-cycle 16x16x4 step.  With blocks as the
    // row quarters t = 16m + 4b + i, its A operand is lane for lane the 16x16x4
    // one; the B column of small tile k is 16 NTF + 4k + l%4, and its result
    // D[t][q] sits in lane 16i + 4b + j.
    constexpr int QREM = QGA % 16;
    constexpr bool SMALL = QREM != 0 && QREM <= 8;
    constexpr int NTF = QGA / 16 + (SMALL || QREM == 0 ? 0 : 1); // 16-column tiles
    constexpr int NSM = SMALL ? (QREM + 3) / 4 : 0;               // 4-column tiles
    constexpr int NCG = NTF + NSM;                                // column groups of the B operand
    static_assert(NTL == NTF + (SMALL ? 1 : 0), "column tiles");
    int bi1[NCG], bi2[NCG];
#pragma unroll
    for (int n = 0; n < NCG; n++) {
        const int q = n >= NTF ? 16 * NTF + 4 * (n - NTF) + (lane & 3) : 16 * n + (lane & 15);
        if (q < QG) {
            const int f = q / nh, c = q % nh;
            bi1[n] = c;
            bi2[n] = f == 0 ? nh + 3 : nh + f - 1;
        } else if (aux && q < QGA) {
            const int x = q - QG;
            bi1[n] = NSRC + (x < NF ? x : NF + (x - NF) / 3);
            bi2[n] = x < NF ? nh + 3 : nh + (x - NF) % 3;
        } else {
            bi1[n] = nh + 3;
            bi2[n] = nh + 3;
        }
    }
    mf_d4 acc[MT][NTF];
    double accs[MT][NSM > 0 ? NSM : 1]; // the 4x4x4 tiles
    auto zero_acc = [&]() {
#pragma unroll
        for (int m = 0; m < MT; m++) {
#pragma unroll
            for (int k = 0; k < NSM; k++) accs[m][k] = 0.0;
#pragma unroll
            for (int n = 0; n < NTF; n++) acc[m][n] = mf_d4{0.0, 0.0, 0.0, 0.0};
        }
    };
    zero_acc();
    const bool deposits = s <= 3;
    // The accumulators' deposit to the grid, for the cell whose box offset
    // along x is w (0, or 1 for a wave's second cell).  For the f64 16x16x4
    // product, element r of lane l holds D[16m + 4r + l/16][16n + l%16] (rows
    // interleaved across the four 16-lane groups, unlike the f32 form's
    // 4(l/16) + r).
    auto flush_direct = [&](int w) {
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int t = 16 * m + 4 * r + (lane >> 4);
                if (t >= NT) continue;
                const int64_t tc = ctab[box_of(t) + w];
#pragma unroll
                for (int n = 0; n < NTF; n++) {
                    const int q = 16 * n + (lane & 15);
                    if (q < QG) unsafeAtomicAdd(&a.gout[tc * QG + q], acc[m][n][r]);
                    else if (aux && q < QGA) unsafeAtomicAdd(&a.aux[tc * NAUX + (q - QG)], acc[m][n][r]);
                }
            }
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int k = 0; k < NSM; k++) {
                const int t = 16 * m + 4 * ((lane >> 2) & 3) + (lane >> 4), q = 16 * NTF + 4 * k + (lane & 3);
                if (t >= NT || q >= QGA) continue;
                const int64_t tc = ctab[box_of(t) + w];
                if (q < QG) unsafeAtomicAdd(&a.gout[tc * QG + q], accs[m][k]);
                else unsafeAtomicAdd(&a.aux[tc * NAUX + (q - QG)], accs[m][k]);
            }
    };

    // The wave's particles as a list of slots: the first cell's in 16-particle
    // tiles (its last tile padded), then the second cell's (CPW = 2), so that a
    // tile never mixes cells; 64 slots (four tiles) per chunk.  With one cell
    // this is the cell's run, 64 particles at a time.
    const int n0 = end - beg, n1 = end1 - beg1;
    const int T0 = (n0 + 15) >> 4;        // tiles of the first cell
    const int NSL = 16 * T0 + n1;         // slots in use
    const int bany = n0 > 0 ? beg : beg1; // a particle of the wave, for inactive lanes' loads
    // (one cell: the loop runs over the cell's run itself, base = beg + sbase)
    const int lo = CPW == 1 ? beg : 0, hi = CPW == 1 ? end : NSL;
    for (int base = lo; base < hi; base += 64) {
        const int sbase = base - lo;
        int64_t ip;
        bool act;
        int pcell;
        if constexpr (CPW == 1) {
            ip = base + lane;
            act = ip < end;
            pcell = cell;
        } else {
            const int sl = sbase + lane;
            const bool first = sl < 16 * T0;
            ip = first ? beg + sl : beg1 + (sl - 16 * T0);
            act = first ? sl < n0 : sl < NSL;
            pcell = first ? cell : cell1; // this lane's particle's cell
        }
        const int64_t ipl = act ? ip : (CPW == 1 ? beg : bany); // a valid index for inactive lanes' loads
        double pos[3];
#pragma unroll
        for (int d = 0; d < 3; d++) pos[d] = a.cur.pos[d][ipl];
        const T pupt = a.cur.pup[3][ipl];
        const T pupv[3] = {a.cur.pup[0][ipl], a.cur.pup[1][ipl], a.cur.pup[2][ipl]};
        // the direction and 1/pupt for the fields (within an ulp of the quotients;
        // the position update keeps the exact quotient)
        const T inv_pupt = T(1) / pupt;
        const T vhat[3] = {pupv[0] * inv_pupt, pupv[1] * inv_pupt, pupv[2] * inv_pupt};
        const T N = a.cur.Nn[0][ipl], Nb = a.cur.Nn[1][ipl];
        int idx[3];
        double wd[3][3];
#pragma unroll
        for (int d = 0; d < 3; d++) shape3((pos[d] - g.plo[d]) * g.dxi[d], g.order[d], idx[d], wd[d]);
        // shape factors and directions of the chunk, for the matrix-core operands
        // of the gather and the deposit (positions are frozen within a step)
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int e = 0; e < 3; e++) Wf[lane * 9 + 3 * d + e] = act ? wd[d][e] : 0.0;
#pragma unroll
        for (int d = 0; d < 3; d++) Src[lane * NSR + nh + d] = act ? vhat[d] : T(0);
        Src[lane * NSR + nh + 3] = T(1);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // the stage state: y_n (stages 0, 1), Y2 and Y4 in Y, Y3 in the
        // accumulator's array (stage 2 writes it there: see below)
        T Y[2 * nh];
        const T* ysrc = s <= 1 ? a.cur.f : (s == 3 ? a.acc : a.Y);
#pragma unroll
        for (int c = 0; c < 2 * nh; c++) Y[c] = ysrc[(int64_t)c * a.np + ipl];
        T Dep[2 * nh]; // state deposited by this launch
        if constexpr (s == 0) {
#pragma unroll
            for (int c = 0; c < 2 * nh; c++) Dep[c] = Y[c];
        } else {
            // ---- gather on the matrix cores, 16 particles per tile
            T inside[nh];
            {
                using MF = Mfma<T>;
                const int g4 = lane >> 4, c16 = lane & 15;
                // B columns past NH only feed result columns that are dropped, so
                // every lane reads an in-range entry, unmasked
                const int cq = g4 * nh + (c16 < nh ? c16 : nh - 1);
                constexpr int NCB = (nh + 3) / 4; // 4-column blocks (FP64 form)
                int cqb[NCB];
#pragma unroll
                for (int cb = 0; cb < NCB; cb++) {
                    const int c = 4 * cb + (lane & 3);
                    cqb[cb] = g4 * nh + (c < nh ? c : nh - 1);
                }
                const int ntile = CPW == 1 ? (min(64, end - base) + 15) >> 4 : min(4, (NSL + 15 - sbase) >> 4);
#pragma unroll 1
                for (int m = 0; m < ntile; m++) {
                    const int p = 16 * m + c16; // the particle of this lane's A rows
                    // the tile's cell's neighbourhood: the second cell's box is one
                    // column further along x (box rows of QG values)
                    const T* Gt = Gs + (CPW == 2 && (sbase >> 4) + m >= T0 ? QG : 0);
                    double w3[9];
#pragma unroll
                    for (int e = 0; e < 9; e++) w3[e] = Wf[p * 9 + e];
                    const T* vp = Src + p * NSR + nh;
                    const T v0 = vp[0], v1 = vp[1], v2 = vp[2];
                    const T sg = g4 == 0 ? T(1) : (g4 == 1 ? -v0 : (g4 == 2 ? -v1 : -v2));
                    // A = (w_x s_g) (w_y w_z): 3 + WY WZ products, then one per target
                    // (FP32-bulk forms w_y w_z in FP32 from the rounded factors, one
                    // conversion per factor instead of one per product: the A
                    // operand is FP32 either way)
                    T ax[WX], ayz[WY * WZ];
#pragma unroll
                    for (int ii = 0; ii < WX; ii++) ax[ii] = T(w3[WX == 1 ? 1 : ii]) * sg;
#pragma unroll
                    for (int kk = 0; kk < WZ; kk++)
#pragma unroll
                        for (int jj = 0; jj < WY; jj++)
                            ayz[kk * WY + jj] = sizeof(T) == 4
                                                     ? T(w3[3 + (WY == 1 ? 1 : jj)]) * T(w3[6 + (WZ == 1 ? 1 : kk)])
                                                     : T(w3[3 + (WY == 1 ? 1 : jj)] * w3[6 + (WZ == 1 ? 1 : kk)]);
                    if constexpr (sizeof(T) == 8) {
                        // FP64: the NH result columns in blocks of 4 on the 4-block
                        // 4x4x4 form (16 cycles each): 3 blocks instead of one 16-column
                        // 16x16x4 (64 cycles) at NF = 3.  Block b is particles
                        // 16m + 4b .. +3; A[b][i][k] sits in lane 16k + 4b + i, which is
                        // the same particle and group as the 16x16x4 operand above;
                        // B[b][k][j] in lane 16k + 4b + j; D[b][i][j] in lane 16i + 4b + j.
                        double accb[NCB];
#pragma unroll
                        for (int cb = 0; cb < NCB; cb++) accb[cb] = 0.0;
#pragma unroll
                        for (int t = 0; t < NT; t++) {
                            const int ii = t % WX, jj = (t / WX) % WY, kk = t / (WX * WY);
                            const T A = ax[ii] * ayz[kk * WY + jj];
#pragma unroll
                            for (int cb = 0; cb < NCB; cb++)
                                accb[cb] = __builtin_amdgcn_mfma_f64_4x4x4f64(A, Gt[box_of(t) * QG + cqb[cb]], accb[cb], 0, 0, 0);
                        }
                        const int pr = 16 * m + 4 * ((lane >> 2) & 3) + (lane >> 4);
#pragma unroll
                        for (int cb = 0; cb < NCB; cb++) {
                            const int c = 4 * cb + (lane & 3);
                            if (c < nh) Src[pr * NSR + c] = T(accb[cb]);
                        }
                    } else {
                        typename MF::acc_t accg = {T(0), T(0), T(0), T(0)};
#pragma unroll
                        for (int t = 0; t < NT; t++) {
                            const int ii = t % WX, jj = (t / WX) % WY, kk = t / (WX * WY);
                            const T A = ax[ii] * ayz[kk * WY + jj];
                            const T B = Gt[box_of(t) * QG + cq];
                            accg = MF::step(A, B, accg);
                        }
                        if (c16 < nh) {
#pragma unroll
                            for (int r = 0; r < 4; r++) Src[(16 * m + MF::row(r, lane)) * NSR + c16] = accg[r];
                        }
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
                for (int c = 0; c < nh; c++) inside[c] = Src[lane * NSR + c];
            }
            // ---- the RK registers this stage reads, as one batch of loads whose
            // latency the Hamiltonian and commutator arithmetic covers
            T yn[2 * nh], ao[2 * nh];
            // The accumulators after stages 1 and 2 are not stored.  With
            // Y2 = y_n + dt/2 k1 and Y3 = y_n + dt/2 k2,
            //   acc1 = y_n + dt/6 k1 = y_n + (Y2 - y_n)/3,
            //   acc2 = acc1 + dt/3 k2 = Y3 + (Y2 - Y3)/3 (y_n cancels; in this
            //   form the rounded 1/3 only scales the small difference, where
            //   Y2/3 + 2 Y3/3 would scale the state by (1/3 + 2/3 rounded) - 1,
            //   3e-8 in FP32, every step),
            // which stages 2 and 3 rebuild from arrays they read: stage 2 keeps
            // Y2 and writes Y3 into the accumulator's array (free until stage 3
            // writes acc3 there), so stage 3 reads y_n, Y2 and Y3.  That saves
            // two writes and a read of the flavour state per step; the change
            // is at rounding level, with multiplies by the rounded 1/3 and 2/3.
            if constexpr (s == 1) {
#pragma unroll
                for (int c = 0; c < 2 * nh; c++) yn[c] = ao[c] = Y[c];
            } else if constexpr (s == 2) {
#pragma unroll
                for (int c = 0; c < 2 * nh; c++) yn[c] = a.cur.f[(int64_t)c * a.np + ipl];
#pragma unroll
                for (int c = 0; c < 2 * nh; c++) ao[c] = yn[c] + (Y[c] - yn[c]) * T(1.0 / 3.0);
            } else if constexpr (s == 3) {
#pragma unroll
                for (int c = 0; c < 2 * nh; c++) yn[c] = a.cur.f[(int64_t)c * a.np + ipl];
#pragma unroll
                for (int c = 0; c < 2 * nh; c++) ao[c] = a.Y[(int64_t)c * a.np + ipl];
#pragma unroll
                for (int c = 0; c < 2 * nh; c++) ao[c] = (ao[c] - Y[c]) * T(1.0 / 3.0) + Y[c];
            } else {
#pragma unroll
                for (int c = 0; c < 2 * nh; c++) yn[c] = s < 4 ? a.cur.f[(int64_t)c * a.np + ipl] : T(0);
#pragma unroll
                for (int c = 0; c < 2 * nh; c++) ao[c] = a.acc[(int64_t)c * a.np + ipl];
            }
            // ---- Hamiltonians / hbar: H = H_vac + coef inside, Hbar = conj(H_vac - coef inside)
            T H[nh], Hb[nh];
#pragma unroll
            for (int c = 0; c < nh; c++) {
                const T hv = T(a.vac[c]) * inv_pupt;
                const T hs = T(a.coef) * inside[c];
                H[c] = hv + hs;
                Hb[c] = is_im_comp<NF>(c) ? hs - hv : hv - hs;
            }
            T k[2 * nh];
            commutator_rhs<NF, T>(H, Y, k, T(1));
            commutator_rhs<NF, T>(Hb, Y + nh, k + nh, T(1));
            const T a_acc = T(dt * bw[s - 1]);
            if constexpr (s < 4) {
                const T a_next = T(dt * cnext[s - 1]);
#pragma unroll
                for (int c = 0; c < 2 * nh; c++) {
                    Dep[c] = a_next * k[c] + yn[c];
                    ao[c] = a_acc * k[c] + ao[c];
                }
                if (act) {
                    // stage 1: Y2 -> Y; stage 2: Y3 -> acc's array; stage 3: acc3 ->
                    // acc, Y4 -> Y (each lane overwrites only the slot it read)
#pragma unroll
                    for (int c = 0; c < 2 * nh; c++) {
                        if (s == 3) a.acc[(int64_t)c * a.np + ip] = ao[c];
                        if (s == 2) a.acc[(int64_t)c * a.np + ip] = Dep[c];
                        else a.Y[(int64_t)c * a.np + ip] = Dep[c];
                    }
                }
                if (s == 1 && act) { // the end-of-step cell, for the sort before stage 4
                    double v[3], npos[3];
#pragma unroll
                    for (int d = 0; d < 3; d++) v[d] = double(pupv[d]) / double(pupt) * phys::c;
                    int nc;
                    end_position(pos, v, dt, g, npos, nc);
                    a.newcell[ip] = nc;
                }
            } else if (act) {
                // y_{n+1} = acc + (dt/6) k4, then the post-step hook (main.cpp:155-199):
                // periodic shift, Renormalize in FP64, and the move to the sorted slot
                double y[2 * nh];
#pragma unroll
                for (int c = 0; c < 2 * nh; c++) y[c] = double(a_acc * k[c] + ao[c]);
                const T L = a.cur.Ll[0][ip], Lb = a.cur.Ll[1][ip];
                const int64_t id = a.cur.id[ip];
                int err = renorm_fp64<NF>(y, double(L), a.maxError);
                if (err == EMU_OK) err = renorm_fp64<NF>(y + nh, double(Lb), a.maxError);
                if (err != EMU_OK) atomicMin(a.err, status_word(id, err));
                double v[3], npos[3];
#pragma unroll
                for (int d = 0; d < 3; d++) v[d] = double(pupv[d]) / double(pupt) * phys::c;
                int nc;
                end_position(pos, v, dt, g, npos, nc);
                const int64_t j = (int64_t)a.mbase[(int64_t)pcell * 27 + a.mv[ip]] + a.rank[ip];
                if (j >= 0 && j < a.np) {
#pragma unroll
                    for (int d = 0; d < 3; d++) a.dst.pos[d][j] = npos[d];
#pragma unroll
                    for (int d = 0; d < 3; d++) a.dst.pup[d][j] = pupv[d];
                    a.dst.pup[3][j] = pupt;
                    a.dst.Nn[0][j] = N;
                    a.dst.Nn[1][j] = Nb;
                    a.dst.Ll[0][j] = L;
                    a.dst.Ll[1][j] = Lb;
                    a.dst.id[j] = id;
#pragma unroll
                    for (int c = 0; c < 2 * nh; c++) a.dst.f[(int64_t)c * a.np + j] = T(y[c]);
                } else {
                    atomicMin(a.err, status_word(id, EMU_ERR_MOVE));
                }
            }
        }
        if (!deposits) continue;
        // ---- stage the chunk's deposit sources in LDS (zeros for inactive lanes)
#pragma unroll
        for (int c = 0; c < nh; c++) {
            const T qc = is_im_comp<NF>(c) ? N * Dep[c] + Nb * Dep[nh + c] : N * Dep[c] - Nb * Dep[nh + c];
            Src[lane * NSR + c] = act ? qc : T(0);
        }
        if (aux) {
#pragma unroll
            for (int x = 0; x < NF; x++)
                Src[lane * NSR + NSRC + x] = act ? N * Dep[herm_re<NF>(x, x)] + Nb * Dep[nh + herm_re<NF>(x, x)] : T(0);
            if constexpr (AUX == 2) {
                int o = NF;
#pragma unroll
                for (int x = 0; x < NF; x++)
#pragma unroll
                    for (int y2 = x + 1; y2 < NF; y2++)
                        Src[lane * NSR + NSRC + (o++)] = act ? Nb * Dep[nh + herm_im<NF>(x, y2)] : T(0);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // ---- D += W^T R, four particles per MFMA
        // KU k-steps (4 KU particles) per iteration, so the operand reads of one
        // overlap the matrix work of the others: two, and four in FP32-bulk
        // stages 1 to 3 (one tile of 16 slots an iteration, so the second
        // cell's tiles still begin at an iteration; FP64 measured no faster);
        // rows past the chunk's particles hold zero sources (inactive lanes), so
        // rounding the count up is exact.  The deposit-only launch keeps one
        // k-step per iteration: there the registers of the second cost a wave of
        // occupancy, and measured slower.  Each operand has one per-lane LDS
        // pointer that advances by a whole iteration; the later k-steps are
        // immediate offsets.  B is the product in T: in FP32-bulk the flux
        // sources q v are FP32 values like q itself, and the sum over particles
        // stays FP64.
        constexpr int KU = s == 0 ? 1 : (sizeof(T) == 4 ? 4 : 2);
        // (32-bit LDS pointers made opaque once, so that the wave's base is not
        // split off again and re-added to every address in the loop)
        typedef __attribute__((address_space(3))) const double* lds_d;
        typedef __attribute__((address_space(3))) const T* lds_t;
        // k-step iterations (4 KU slots each) up to the chunk's last particle.
        // With two cells, the first cell is complete at the iteration where the
        // second cell's tiles begin (kb, at a tile boundary): its deposit is
        // flushed there and the accumulators start again from zero.
        const int kend = ((CPW == 1 ? min(64, end - base) : min(NSL - sbase, 64)) + 4 * KU - 1) / (4 * KU);
        const int kb = CPW == 2 && n0 > 0 && sbase < n0 && n0 <= sbase + 64 ? (16 * T0 - sbase) / (4 * KU) : -1;
        lds_d wa[MT][3];
        lds_t sb[NCG][2];
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                wa[m][d] = (lds_d)(Wf + (lane >> 4) * 9 + tdx[m][d]);
                asm volatile("" : "+v"(wa[m][d]));
            }
#pragma unroll
        for (int n = 0; n < NCG; n++) {
            sb[n][0] = (lds_t)(Src + (lane >> 4) * NSR + bi1[n]);
            sb[n][1] = (lds_t)(Src + (lane >> 4) * NSR + bi2[n]);
            asm volatile("" : "+v"(sb[n][0]));
            asm volatile("" : "+v"(sb[n][1]));
        }
#pragma unroll 1
        for (int kp = 0; kp < kend; kp++) {
            if (CPW == 2 && kp == kb) {
                flush_direct(0);
                zero_acc();
            }
#pragma unroll
            for (int h = 0; h < KU; h++) {
                double A[MT];
#pragma unroll
                for (int m = 0; m < MT; m++) A[m] = wa[m][0][36 * h] * wa[m][1][36 * h] * wa[m][2][36 * h];
#pragma unroll
                for (int n = 0; n < NCG; n++) {
                    const double B = double(sb[n][0][4 * NSR * h] * sb[n][1][4 * NSR * h]);
                    if (n >= NTF) {
#pragma unroll
                        for (int m = 0; m < MT; m++)
                            accs[m][n - NTF] = __builtin_amdgcn_mfma_f64_4x4x4f64(A[m], B, accs[m][n - NTF], 0, 0, 0);
                    } else {
#pragma unroll
                        for (int m = 0; m < MT; m++)
                            acc[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[m], B, acc[m][n], 0, 0, 0);
                    }
                }
            }
#pragma unroll
            for (int m = 0; m < MT; m++)
#pragma unroll
                for (int d = 0; d < 3; d++) wa[m][d] += 36 * KU;
#pragma unroll
            for (int n = 0; n < NCG; n++) {
                sb[n][0] += 4 * NSR * KU;
                sb[n][1] += 4 * NSR * KU;
            }
        }
        if (CPW == 2 && kb == kend) {
            flush_direct(0);
            zero_acc();
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (!deposits) return;
    if constexpr (!MERGE) {
        // ---- flush this wave's (last) cell's deposit straight to the grid
        if constexpr (CPW == 2) {
            if (n1 > 0) flush_direct(1);
            return;
        }
        if (end <= beg) return;
        flush_direct(0);
        return;
    }
    // ---- merge the waves' deposits over the box in LDS, then flush
    double* Mb = lw; // [NB][QGA], over the staged box and the waves' regions (not the cell table)
    __syncthreads();
    for (int e = threadIdx.x; e < NB * QGA; e += 64 * WPB) Mb[e] = 0.0;
    __syncthreads();
    if (end > beg) {
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int t = 16 * m + 4 * r + (lane >> 4);
                if (t >= NT) continue;
                const int bi = box_of(t);
#pragma unroll
                for (int n = 0; n < NTF; n++) {
                    const int q = 16 * n + (lane & 15);
                    if (q < QGA) unsafeAtomicAdd(&Mb[bi * QGA + q], acc[m][n][r]);
                }
            }
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int k = 0; k < NSM; k++) {
                const int t = 16 * m + 4 * ((lane >> 2) & 3) + (lane >> 4), q = 16 * NTF + 4 * k + (lane & 3);
                if (t < NT && q < QGA) unsafeAtomicAdd(&Mb[box_of(t) * QGA + q], accs[m][k]);
            }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < NB * QGA; e += 64 * WPB) {
        const double v = Mb[e];
        if (v == 0.0) continue;
        const int bi = e / QGA, q = e - bi * QGA;
        const int64_t tc = ctab[bi];
        if (q < QG) unsafeAtomicAdd(&a.gout[tc * QG + q], v);
        else unsafeAtomicAdd(&a.aux[tc * NAUX + (q - QG)], v);
    }
}

} // namespace emu
