// emu_common.h -- constants, Hermitian component algebra and the particle
// shape factors shared by every kernel and by the host code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define EMU_HD __host__ __device__ __forceinline__

namespace emu {

// Constants.H:6-26, evaluated in the same operation order as the constexprs.
namespace phys {
constexpr double eV = 1.60218e-12;  // erg
constexpr double c = 2.99792458e10; // cm/s
constexpr double c2 = c * c;
constexpr double c4 = c2 * c2;
constexpr double hbar = 1.05457266e-27; // erg s
constexpr double hbarc = hbar * c;
constexpr double GF = 1.1663787e-5 / (1e9 * 1e9 * eV * eV) * hbarc * hbarc * hbarc; // erg cm^3
constexpr double Mp = 1.6726219e-24; // g
} // namespace phys

// Number of real components of an NF x NF Hermitian matrix.
template <int NF> constexpr int NH = NF * NF;

// Index of the real part of X_ij (i <= j after ordering) in the component
// order of Scripts/symbolic_hermitians/HermitianUtils.py:77-83,170-177:
// for i: for j >= i: (i,j,Re) then (i,j,Im) when j > i.
template <int NF> EMU_HD constexpr int herm_re(int i, int j)
{
    const int a = i < j ? i : j, b = i < j ? j : i;
    int n = 0;
    for (int r = 0; r < a; ++r) n += 1 + 2 * (NF - 1 - r);
    return n + (b == a ? 0 : 1 + 2 * (b - a - 1));
}
template <int NF> EMU_HD constexpr int herm_im(int i, int j) { return herm_re<NF>(i, j) + 1; }

// Particle attribute order (PIdx, FlavoredNeutrinoContainer.H:12-26 with the
// generated enum, generate_code.py:93-104).
template <int NF> struct PIdx {
    enum : int {
        time = 0, x, y, z, pupx, pupy, pupz, pupt, N, L, f,
        Nbar = f + NH<NF>, Lbar, fbar, nattribs = fbar + NH<NF>
    };
};

// Grid component order (GIdx, Evolve.H:11-27 with generate_code.py:127-135).
template <int NF> struct GIdx {
    enum : int { rho = 0, T, Ye, dep0 = 3, ncomp = 3 + 8 * NH<NF> };
    // v: 0 = N, 1 = Fx, 2 = Fy, 3 = Fz; bar: 0 = nu, 1 = nubar
    static EMU_HD constexpr int var(int v, int bar) { return 3 + (2 * v + bar) * NH<NF>; }
};

// ---------------------------------------------------------------------------
// Shape factors, Source/ParticleInterpolator.H:34-131.  Cell index and offset
// follow the reference's truncation rules exactly (the index is the binning
// contract); contraction is disabled so the weights round as written.
// ---------------------------------------------------------------------------
EMU_HD int nearest_cell_center_index(double delta)
{
#pragma clang fp contract(off)
    const double dstar = delta - 0.5 < 0.0 ? delta - 1.0 : delta;
    return static_cast<int>(dstar);
}

EMU_HD double nearest_cell_center_offset(double delta)
{
#pragma clang fp contract(off)
    const double fraction = delta - static_cast<int>(delta);
    return delta <= 0.0 ? fraction + 0.5 : fraction - 0.5;
}

// Order-generic weights over a 3-wide window [idx-1, idx, idx+1] relative to
// the particle's cell idx; order 1 leaves one edge weight zero and order 0
// uses only the middle entry.  first = idx - 1 in every case.
EMU_HD void shape3(double delta, int order, int& idx, double w[3])
{
#pragma clang fp contract(off)
    idx = nearest_cell_center_index(delta);
    const double o = nearest_cell_center_offset(delta);
    if (order > 1) {
        w[0] = 0.5 * (0.5 - o) * (0.5 - o);
        w[1] = 0.75 - o * o;
        w[2] = 0.5 * (0.5 + o) * (0.5 + o);
    } else if (order > 0) {
        if (o >= 0) { w[0] = 0.0; w[1] = 1.0 - o; w[2] = o; }
        else { w[0] = -o; w[1] = 1.0 + o; w[2] = 0.0; }
    } else {
        w[0] = 0.0; w[1] = 1.0; w[2] = 0.0;
    }
}

EMU_HD int pmod(int i, int n)
{
    const int r = i % n;
    return r < 0 ? r + n : r;
}

// ---------------------------------------------------------------------------
// dfdt = (H f - f H)(-i/hbar) on real components (generate_code.py:363-389).
// With A = H f, C = A - A^dagger, so only A is formed.  'scale' multiplies
// the result: 1/hbar for the reference form, 1 when H is pre-divided by hbar.
// ---------------------------------------------------------------------------
template <int NF, typename T>
EMU_HD void commutator_rhs(const T* H, const T* f, T* out, T scale)
{
    // C = A - A^dagger with A = H f needs Im A_ii and A_ij, A_ji for i < j only.
    // Each is one FMA chain; the products with a zero factor (the imaginary
    // parts of diagonal entries) are left out at compile time, and the chains
    // start from -0, which leaves the first product exact.
    auto re = [&](const T* X, int i, int j) -> T { return X[herm_re<NF>(i < j ? i : j, i < j ? j : i)]; };
    auto im = [&](const T* X, int i, int j) -> T { // i != j
        return i < j ? X[herm_im<NF>(i, j)] : -X[herm_im<NF>(j, i)];
    };
    auto a_re = [&](int i, int j) -> T {
        T s = T(-0.0);
#pragma unroll
        for (int k = 0; k < NF; k++) {
            s = fma(re(H, i, k), re(f, k, j), s);
            if (i != k && k != j) s = fma(-im(H, i, k), im(f, k, j), s);
        }
        return s;
    };
    auto a_im = [&](int i, int j) -> T {
        T s = T(-0.0);
#pragma unroll
        for (int k = 0; k < NF; k++) {
            if (k != j) s = fma(re(H, i, k), im(f, k, j), s);
            if (i != k) s = fma(im(H, i, k), re(f, k, j), s);
        }
        return s;
    };
#pragma unroll
    for (int i = 0; i < NF; i++) {
        // C_ii = 2i Im A_ii; (-i) C_ii = 2 Im A_ii
        out[herm_re<NF>(i, i)] = T(2) * a_im(i, i) * scale;
#pragma unroll
        for (int j = i + 1; j < NF; j++) {
            // C_ij = A_ij - conj(A_ji); (-i)(Cr + i Ci) = Ci - i Cr
            const T cr = a_re(i, j) - a_re(j, i);
            const T ci = a_im(i, j) + a_im(j, i);
            out[herm_re<NF>(i, j)] = ci * scale;
            out[herm_im<NF>(i, j)] = -cr * scale;
        }
    }
}

// |X|_SU^2, HermitianUtils.py:102-120.
template <int NF, typename T> EMU_HD T su_mag2(const T* h)
{
    T m2 = T(0);
#pragma unroll
    for (int i = 0; i < NF; i++)
#pragma unroll
        for (int j = i + 1; j < NF; j++) {
            const T re = h[herm_re<NF>(i, j)], im = h[herm_im<NF>(i, j)];
            m2 += re * re + im * im;
        }
#pragma unroll
    for (int l = 1; l < NF; l++) {
        T bc = T(0);
#pragma unroll
        for (int i = 1; i <= l; i++) bc += h[herm_re<NF>(i - 1, i - 1)];
        bc -= T(l) * h[herm_re<NF>(l, l)];
        bc *= T(l == 1 ? 1.0 : 0.57735026918962573); // sqrt(2/(l(l+1))) for l = 1, 2
        m2 += (bc / T(2)) * (bc / T(2));
    }
    return m2;
}

} // namespace emu
