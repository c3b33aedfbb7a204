"""Pins the FP64 CPU oracle before it is trusted as the checker.

Each test checks one oracle function against something the oracle's own
formula does not contain: the reference's golden notebook output, closed
forms, brute force with an independent formulation, invariants, or the
reference's analytic acceptance tests.
"""
import json
import math
import os

import numpy as np
import pytest

import cases
import kat

NB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "notebook_commutator_nf3.json")


def herm_index(nf):
    names = []
    for i in range(nf):
        for j in range(i, nf):
            names.append(f"{i}{j}_Re")
            if j > i:
                names.append(f"{i}{j}_Im")
    return {n: k for k, n in enumerate(names)}


def to_complex(h, nf):
    idx = herm_index(nf)
    M = np.zeros((nf, nf), dtype=complex)
    for i in range(nf):
        for j in range(i, nf):
            v = h[idx[f"{i}{j}_Re"]] + (1j * h[idx[f"{i}{j}_Im"]] if j > i else 0)
            M[i, j] = v
            M[j, i] = np.conj(v)
    return M


def from_complex(M, nf):
    idx = herm_index(nf)
    h = np.zeros(nf * nf)
    for i in range(nf):
        for j in range(i, nf):
            h[idx[f"{i}{j}_Re"]] = M[i, j].real
            if j > i:
                h[idx[f"{i}{j}_Im"]] = M[i, j].imag
    return h


# --------------------------------------------------------------------------
# commutator: golden notebook output, 2x2 closed form, invariants
# --------------------------------------------------------------------------
def test_commutator_matches_notebook_golden(oracle):
    table = json.load(open(NB))["terms"]
    idx = herm_index(3)
    rng = np.random.default_rng(7)
    for _ in range(20):
        H = rng.normal(size=9)
        f = rng.normal(size=9)
        got = oracle.commutator_rhs(3, H, f)
        for comp, terms in table.items():
            D = sum(c * H[idx[a]] * f[idx[b]] for c, a, b in terms)
            want = -D / oracle.hbar  # dfdt = -(i/hbar)[H,f] = -(1/hbar) * i[H,f]
            assert got[idx[comp]] == pytest.approx(want, rel=1e-12, abs=1e-12 * np.abs(got).max())


def test_commutator_nf2_closed_form_and_embedding(oracle):
    rng = np.random.default_rng(8)
    for _ in range(10):
        H = rng.normal(size=4)
        f = rng.normal(size=4)
        d2 = oracle.commutator_rhs(2, H, f)
        # dfdt00 = 2 (V01_Im f01_Re - V01_Re f01_Im) / hbar
        assert d2[0] == pytest.approx(2 * (H[2] * f[1] - H[1] * f[2]) / oracle.hbar, rel=1e-13)
        # the 2x2 block of the 3x3 commutator with a decoupled third flavour
        H3 = from_complex(np.pad(to_complex(H, 2), ((0, 1), (0, 1))), 3)
        f3 = from_complex(np.pad(to_complex(f, 2), ((0, 1), (0, 1))), 3)
        d3 = oracle.commutator_rhs(3, H3, f3)
        np.testing.assert_allclose(to_complex(d3, 3)[:2, :2], to_complex(d2, 2), rtol=1e-13,
                                   atol=1e-13 * np.abs(d2).max())


@pytest.mark.parametrize("nf", [2, 3])
def test_commutator_traceless_and_identity_shift(oracle, nf):
    rng = np.random.default_rng(9)
    H = rng.normal(size=nf * nf)
    f = rng.normal(size=nf * nf)
    d = to_complex(oracle.commutator_rhs(nf, H, f), nf)
    assert abs(np.trace(d)) < 1e-12 * np.abs(d).max()
    Hs = from_complex(to_complex(H, nf) + 3.7 * np.eye(nf), nf)
    np.testing.assert_allclose(oracle.commutator_rhs(nf, Hs, f), oracle.commutator_rhs(nf, H, f),
                               rtol=1e-12, atol=1e-12 * np.abs(d).max())


# --------------------------------------------------------------------------
# SU(n) magnitude and the mass matrix
# --------------------------------------------------------------------------
def test_su_magnitude_closed_forms(oracle):
    assert oracle.su_mag(2, [1, 0, 0, 0]) == pytest.approx(0.5, rel=1e-15)
    assert oracle.su_mag(3, [1, 0, 0, 0, 0, 0, 0, 0, 0]) == pytest.approx(math.sqrt(1 / 3), rel=1e-15)
    rng = np.random.default_rng(1)
    for nf in (2, 3):
        h = rng.normal(size=nf * nf)
        shifted = from_complex(to_complex(h, nf) + 2.5 * np.eye(nf), nf)
        assert oracle.su_mag(nf, shifted) == pytest.approx(oracle.su_mag(nf, h), rel=1e-13)
        # |X|_SU^2 = (1/2) tr(X0^2) with X0 the traceless part (Gell-Mann normalisation)
        X = to_complex(h, nf)
        X0 = X - np.trace(X) / nf * np.eye(nf)
        assert oracle.su_mag(nf, h) ** 2 == pytest.approx(0.5 * np.trace(X0 @ X0).real, rel=1e-12)


def _deck(oracle, **kw):
    d = oracle.Deck(**kw)
    return d, d.params()


def test_mass_matrix_nf2_closed_form(oracle):
    d, p = _deck(oracle, nf=2, mass1_eV=-0.008596511, mass2_eV=0.0, theta12_degrees=33.82)
    M = oracle.mass_matrix(2, p)
    m1, m2 = p.mass[0] ** 2, p.mass[1] ** 2
    c, s = math.cos(p.theta12), math.sin(p.theta12)
    np.testing.assert_allclose(M, [c * c * m1 + s * s * m2, c * s * (m2 - m1), 0.0, s * s * m1 + c * c * m2],
                               rtol=1e-14, atol=1e-14 * m1)


def test_mass_matrix_nf3_quirk_and_spectrum(oracle):
    d, p = _deck(oracle, nf=3, mass1_eV=0.0, mass2_eV=math.sqrt(7.4e-5), mass3_eV=math.sqrt(2.5e-3),
                 theta12_degrees=33.82, theta13_degrees=48.3, theta23_degrees=8.61, deltaCP_degrees=222)
    M = to_complex(oracle.mass_matrix(3, p), 3)
    ev = np.sort(np.linalg.eigvalsh(M))
    np.testing.assert_allclose(ev, np.sort(np.array(p.mass) ** 2), rtol=1e-10, atol=1e-12 * ev.max())
    # theta23 is read but unused (generate_code.py:202-209)
    d2, p2 = _deck(oracle, nf=3, mass1_eV=0.0, mass2_eV=math.sqrt(7.4e-5), mass3_eV=math.sqrt(2.5e-3),
                   theta12_degrees=33.82, theta13_degrees=48.3, theta23_degrees=40.0, deltaCP_degrees=222)
    np.testing.assert_array_equal(oracle.mass_matrix(3, p2), oracle.mass_matrix(3, p))
    # With theta12 = deltaCP = 0, U = U23 U13 is ONE rotation by 2*theta13 in the (0,2)
    # plane (U23 is built from theta13), so M2_00 = cos^2(2 th13) m1^2 + sin^2(2 th13) m3^2.
    d3, p3 = _deck(oracle, nf=3, mass1_eV=0.3, mass2_eV=0.0, mass3_eV=0.05, theta13_degrees=20.0)
    M3 = oracle.mass_matrix(3, p3)
    t = 2 * p3.theta13
    assert M3[0] == pytest.approx(math.cos(t) ** 2 * p3.mass[0] ** 2 + math.sin(t) ** 2 * p3.mass[2] ** 2,
                                  rel=1e-12)


# --------------------------------------------------------------------------
# ParticleInterpolator
# --------------------------------------------------------------------------
def bspline2(x):
    ax = abs(x)
    if ax <= 0.5:
        return 0.75 - ax * ax
    if ax < 1.5:
        return 0.5 * (1.5 - ax) ** 2
    return 0.0


def test_interp_closed_forms(oracle):
    rng = np.random.default_rng(3)
    for delta in np.concatenate([rng.uniform(0, 16, 200), [0.0, 0.5, 1.0, 1.5, 15.999999999, 7.25]]):
        first, w = oracle.interp(float(delta), 2)
        assert w.sum() == pytest.approx(1.0, abs=1e-15)
        # brute force: the continuous quadratic B-spline centred on the particle,
        # evaluated at the three cell centres of the stencil
        for k in range(3):
            assert w[k] == pytest.approx(bspline2(delta - (first + k + 0.5)), abs=2e-16)
        if delta > 0:
            assert first == math.floor(delta) - 1
        f1, w1 = oracle.interp(float(delta), 1)
        assert w1[0] + w1[1] == pytest.approx(1.0, abs=1e-15)
        for k in range(2):  # cloud-in-cell hat function
            assert w1[k] == pytest.approx(max(0.0, 1 - abs(delta - (f1 + k + 0.5))), abs=2e-16)
        f0, w0 = oracle.interp(float(delta), 0)
        assert w0[0] == 1.0 and (delta <= 0 or f0 == math.floor(delta))
    # delta == 0 is treated as cell -1 with offset +1/2 (ParticleInterpolator.H:109-127)
    first, w = oracle.interp(0.0, 2)
    assert first == -2 and list(w) == [0.0, 0.5, 0.5]


# --------------------------------------------------------------------------
# deposit_to_mesh
# --------------------------------------------------------------------------
def random_particles(oracle, nf, ncell, L, n, rng, isotropic_pairs=False):
    d = oracle.Deck(nf=nf, ncell=ncell, Lx=L[0], Ly=L[1], Lz=L[2])
    parts = np.zeros(n, dtype=oracle.particle_dtype(nf))
    K = {nm: i for i, nm in enumerate(oracle.attribute_names(nf))}
    pos = rng.uniform(0, 1, size=(n, 3)) * np.array(L)
    parts["pos"] = pos
    rd = parts["rdata"]
    rd[:, 1:4] = pos
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    E = 50e6 * oracle.eV * rng.uniform(0.5, 2.0, size=n)
    rd[:, K["pupt"]] = E
    rd[:, K["pupx"]], rd[:, K["pupy"]], rd[:, K["pupz"]] = u[:, 0] * E, u[:, 1] * E, u[:, 2] * E
    rd[:, K["N"]] = rng.uniform(1e30, 2e30, n)
    rd[:, K["Nbar"]] = rng.uniform(1e30, 2e30, n)
    nh = nf * nf
    for t in ("", "bar"):
        fs = rng.normal(size=(n, nh)) * 0.1
        k0 = K["f00_Re" + t]
        rd[:, k0:k0 + nh] = fs
    return d, d.params(), parts, K


def bspline_weights(delta, order, n):
    """Independent periodic deposit weights on cell centres: the quadratic
    B-spline (order 2), the cloud-in-cell hat (order 1) or the nearest cell
    (order 0), centred on the particle."""
    w = np.zeros(n)
    for T in range(-3, n + 3):
        x = delta - (T + 0.5)
        if order == 2:
            v = bspline2(x)
        elif order == 1:
            v = max(0.0, 1.0 - abs(x))
        else:
            v = 1.0 if T == math.floor(delta) else 0.0
        w[T % n] += v
    return w


@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("nf", [2, 3])
def test_deposit_brute_force_and_conservation(oracle, nf, order):
    rng = np.random.default_rng(11 + order)
    ncell, L = (4, 3, 5), (1.0, 0.75, 1.25)
    d, p, parts, K = random_particles(oracle, nf, ncell, L, 60, rng)
    d.shape_order = order
    p = d.params()
    g = oracle.Grid(nf, p)
    oracle.deposit_to_mesh(nf, p, parts, g)
    valid = g.valid()  # [comp, k, j, i]
    nh = nf * nf
    want = np.zeros_like(valid)
    dx = np.array(L) / np.array(ncell)
    for q in parts:
        wx = bspline_weights(q["pos"][0] / dx[0], order, ncell[0])
        wy = bspline_weights(q["pos"][1] / dx[1], order, ncell[1])
        wz = bspline_weights(q["pos"][2] / dx[2], order, ncell[2])
        W = wz[:, None, None] * wy[None, :, None] * wx[None, None, :]
        r = q["rdata"]
        vhat = r[K["pupx"]:K["pupx"] + 3] / r[K["pupt"]]
        for b, t in enumerate(("", "bar")):
            Nf = r[K["N" + t]] * r[K["f00_Re" + t]:K["f00_Re" + t] + nh]
            for v in range(4):
                scale = 1.0 if v == 0 else vhat[v - 1]
                c0 = 3 + (2 * v + b) * nh
                want[c0:c0 + nh] += (Nf * scale)[:, None, None, None] * W[None]
    scale = np.abs(want).max(axis=(1, 2, 3), keepdims=True) + 1e-300
    np.testing.assert_allclose(valid[3:] / scale[3:], want[3:] / scale[3:], atol=1e-13)
    # conservation: sum over cells of N_c = sum over particles of N f_c
    for b, t in enumerate(("", "bar")):
        tot = (parts["rdata"][:, K["N" + t]][:, None] * parts["rdata"][:, K["f00_Re" + t]:K["f00_Re" + t] + nh]).sum(0)
        c0 = 3 + b * nh
        np.testing.assert_allclose(valid[c0:c0 + nh].sum(axis=(1, 2, 3)), tot, rtol=1e-12,
                                   atol=1e-12 * np.abs(tot).max())


def test_deposit_isotropic_zero_flux(oracle):
    """uniform_sphere_xyz adds exact antipodes (FlavoredNeutrinoContainerInit.cpp:28-30):
    an isotropic, identically weighted set deposits zero flux in every cell."""
    d = oracle.Deck(nf=2, simulation_type=4, ncell=(4, 4, 4), nphi_equator=8, st4_ndens=4.89e32,
                    st4_ndensbar=4.89e32, st4_fluxfac=0.0, st4_fluxfacbar=0.0, st4_amplitude=0.0)
    parts = oracle.init_particles(d)
    p = d.params()
    g = oracle.Grid(2, p)
    oracle.deposit_to_mesh(2, p, parts, g)
    v = g.valid()
    Nmax = np.abs(v[3:3 + 8]).max()
    assert np.abs(v[3 + 8:]).max() <= 1e-14 * Nmax


# --------------------------------------------------------------------------
# interpolate_rhs_from_mesh: independent matrix-form Hamiltonian
# --------------------------------------------------------------------------
@pytest.mark.parametrize("nf", [2, 3])
def test_interpolate_rhs_matrix_form(oracle, nf):
    rng = np.random.default_rng(12)
    ncell, L = (3, 4, 3), (0.6, 0.8, 0.6)
    d, p, parts, K = random_particles(oracle, nf, ncell, L, 25, rng)
    d.rho_g_ccm, d.Ye = 1e9, 0.4
    d.mass1_eV, d.mass2_eV, d.mass3_eV = 0.0, 0.02, 0.05
    d.theta12_degrees, d.theta13_degrees, d.deltaCP_degrees = 33.0, 9.0, 200.0
    p = d.params()
    g = oracle.Grid(nf, p)
    oracle.deposit_to_mesh(nf, p, parts, g)
    oracle.fill_boundary(nf, p, g)
    rhs = parts.copy()
    oracle.interpolate_rhs_from_mesh(nf, p, rhs, g)
    nh = nf * nf
    valid = g.valid()
    dx = np.array(L) / np.array(ncell)
    M2 = to_complex(oracle.mass_matrix(nf, p), nf)
    vol = dx.prod()
    for q, out in zip(parts, rhs):
        r = q["rdata"]
        E = r[K["pupt"]]
        vhat = r[K["pupx"]:K["pupx"] + 3] / E
        wx = bspline_weights(q["pos"][0] / dx[0], 2, ncell[0])
        wy = bspline_weights(q["pos"][1] / dx[1], 2, ncell[1])
        wz = bspline_weights(q["pos"][2] / dx[2], 2, ncell[2])
        W = wz[:, None, None] * wy[None, :, None] * wx[None, None, :]
        HSI = np.zeros((nf, nf), dtype=complex)
        fields = []
        for b in range(2):
            per_v = []
            for v in range(4):
                c0 = 3 + (2 * v + b) * nh
                per_v.append(to_complex((valid[c0:c0 + nh] * W[None]).sum(axis=(1, 2, 3)), nf))
            fields.append(per_v[0] - vhat[0] * per_v[1] - vhat[1] * per_v[2] - vhat[2] * per_v[3])
        HSI = math.sqrt(2) * oracle.GF / vol * (fields[0] - np.conj(fields[1]))
        Hm = np.zeros((nf, nf), dtype=complex)
        Hm[0, 0] = math.sqrt(2) * oracle.GF * d.rho_g_ccm * d.Ye / oracle.Mp
        Hvac = M2 * oracle.c4 / (2 * E)
        H = Hvac + HSI + Hm
        Hbar = np.conj(Hvac - HSI - Hm)
        for b, t in enumerate(("", "bar")):
            f = to_complex(r[K["f00_Re" + t]:K["f00_Re" + t] + nh], nf)
            Hx = H if b == 0 else Hbar
            dfdt = -1j / oracle.hbar * (Hx @ f - f @ Hx)
            got = to_complex(out["rdata"][K["f00_Re" + t]:K["f00_Re" + t] + nh], nf)
            np.testing.assert_allclose(got, dfdt, rtol=1e-10, atol=1e-11 * np.abs(dfdt).max())
        o = out["rdata"]
        np.testing.assert_allclose(o[K["x"]:K["x"] + 3], vhat * oracle.clight, rtol=1e-15)
        assert o[K["time"]] == 1.0
        assert o[K["N"]] == o[K["Nbar"]] == o[K["L"]] == o[K["Lbar"]] == o[K["pupt"]] == 0.0


# --------------------------------------------------------------------------
# compute_dt, Renormalize, periodic wrap
# --------------------------------------------------------------------------
def test_compute_dt_uniform_closed_form(oracle):
    d = oracle.Deck(nf=3, simulation_type=4, ncell=(4, 4, 4), Lx=2.0, Ly=2.0, Lz=2.0, nphi_equator=8,
                    st4_ndens=4.89e32, st4_ndensbar=4.89e32, st4_fluxfac=0.0, st4_fluxfacbar=0.0,
                    cfl_factor=0.4, flavor_cfl_factor=0.4, mass1_eV=0.0, mass2_eV=0.0086, mass3_eV=0.05)
    parts = oracle.init_particles(d)
    p = d.params()
    g = oracle.Grid(3, p)
    oracle.deposit_to_mesh(3, p, parts, g)
    oracle.fill_boundary(3, p, g)
    vv = oracle.vvac_max(3, p, parts)
    err, dt = oracle.compute_dt(3, p, g, vv)
    assert err == 0
    # every cell holds n*dV neutrinos of flavour e: V_stupid = n dV * 2 NF sqrt2 GF / dV
    Vst = 4.89e32 * 2 * 3 * math.sqrt(2) * oracle.GF
    m2 = (np.array(p.mass) ** 2)
    X0 = np.diag(m2) - m2.mean() * np.eye(3)
    vvac = math.sqrt(0.5 * np.trace(X0 @ X0)) * oracle.c4 / (50e6 * oracle.eV)
    assert vv == pytest.approx(vvac, rel=1e-12)
    dt_fl = oracle.hbar / (Vst + vvac) * 0.4
    dt_tr = 0.5 / oracle.clight * 0.4
    assert dt == pytest.approx(min(dt_fl, dt_tr), rel=1e-12)


def _herm(vals, names, tmpl, nf):
    """Complex NF x NF Hermitian matrix from its stored real components."""
    X = np.zeros((nf, nf), dtype=complex)
    for i in range(nf):
        for j in range(i, nf):
            re = vals[names.index(tmpl.format(i, j, "Re"))]
            im = vals[names.index(tmpl.format(i, j, "Im"))] if j > i else 0.0
            X[i, j] = re + 1j * im
            X[j, i] = re - 1j * im
    return X


def _su2(X):
    """|X|^2 in the SU(N) generator basis = Tr(X0^2)/2, X0 the traceless part."""
    X0 = X - np.trace(X) / X.shape[0] * np.eye(X.shape[0])
    return 0.5 * np.trace(X0 @ X0).real


@pytest.mark.parametrize("nf", [2, 3])
@pytest.mark.parametrize("speedup", [0.0, 1.0, 4.0, 1e9])
def test_compute_dt_adaptive_closed_form(oracle, nf, speedup):
    """compute_dt's two flavor bounds from matrix algebra (Evolve.cpp:36-74,
    generate_code.py:258-294): V_adaptive = |N - conj(Nbar) + matter|_SU and
    |F_d - Fbar_d|_SU (no conjugate on the flux), times sqrt2 GF / dV;
    V_stupid = 2 NF sqrt2 GF / dV times the largest diagonal or matter term;
    max_adaptive_speedup > 1 picks min(speedup dt_stupid, dt_adaptive).  Every
    cell of the fab, ghosts included, is reduced, so one cell carries the
    largest values."""
    d = oracle.Deck(nf=nf, ncell=(3, 2, 2), Lx=1.5, Ly=2.0, Lz=1.0, cfl_factor=-1.0, flavor_cfl_factor=0.3,
                    max_adaptive_speedup=speedup)
    p = d.params()
    g = oracle.Grid(nf, p)
    names = oracle.grid_names(nf)
    rng = np.random.default_rng(7 + nf)
    fab = g.fab()
    vals = rng.uniform(-1.0, 1.0, size=len(names)) * 1e30
    for t in ("", "bar"):
        for i in range(nf):
            vals[names.index("N{0}{0}_Re{1}".format(i, t))] = abs(vals[names.index("N{0}{0}_Re{1}".format(i, t))]) * 3
    vals[names.index("rho")], vals[names.index("T")], vals[names.index("Ye")] = 2e6, 10.0, 0.4
    fab[:] = 0.25 * vals[:, None, None, None]
    fab[:, 0, 1, 2] = vals  # a ghost cell (ng >= 1 with shape order 2)
    dV = (1.5 / 3) * (2.0 / 2) * (1.0 / 2)
    matter = 2e6 * 0.4 / oracle.Mp * dV
    HSI = _herm(vals, names, "N{}{}_{}", nf) - _herm(vals, names, "N{}{}_{}bar", nf).conj()
    HSI[0, 0] += matter
    V2 = _su2(HSI)
    for dd in "xyz":
        V2 += _su2(_herm(vals, names, "F" + dd + "{}{}_{}", nf) - _herm(vals, names, "F" + dd + "{}{}_{}bar", nf))
    Vad = math.sqrt(V2) * math.sqrt(2) * oracle.GF / dV
    diag = [vals[names.index("N{0}{0}_Re{1}".format(i, t))] for i in range(nf) for t in ("", "bar")]
    Vst = max(max(diag), matter) * 2 * nf * math.sqrt(2) * oracle.GF / dV
    vv = 3e-19
    dt_ad = oracle.hbar / (Vad + vv) * 0.3
    dt_st = oracle.hbar / (Vst + vv) * 0.3
    want = min(dt_st * speedup, dt_ad) if speedup > 1 else dt_st
    err, dt = oracle.compute_dt(nf, p, g, vv)
    assert err == 0
    assert dt == pytest.approx(want, rel=1e-12)
    if speedup == 1e9:
        assert dt == pytest.approx(dt_ad, rel=1e-12) and dt_ad > 1.5 * dt_st


def test_compute_dt_error(oracle):
    d = oracle.Deck(nf=2, cfl_factor=0.0, flavor_cfl_factor=0.0)
    p = d.params()
    g = oracle.Grid(2, p)
    err, _ = oracle.compute_dt(2, p, g, 0.0)
    assert err == oracle.ERR_DT


@pytest.mark.parametrize("nf", [2, 3])
def test_renormalize_invariants_and_errors(oracle, nf):
    d = oracle.Deck(nf=nf, simulation_type=4, ncell=(2, 2, 2), nphi_equator=4, st4_ndens=1e32,
                    st4_ndensbar=1e32, st4_amplitude=1e-3, maxError=1e-6)
    parts = oracle.init_particles(d)
    p = d.params()
    K = {nm: i for i, nm in enumerate(oracle.attribute_names(nf))}
    nh = nf * nf
    rng = np.random.default_rng(5)
    for t in ("", "bar"):
        k0 = K["f00_Re" + t]
        parts["rdata"][:, k0:k0 + nh] *= 1 + rng.uniform(-5e-5, 5e-5, size=(len(parts), 1))
    work = parts.copy()
    err, bad = oracle.renormalize(nf, p, work)
    assert err == 0 and bad == -1
    diag = [herm_index(nf)[f"{i}{i}_Re"] for i in range(nf)]
    for t in ("", "bar"):
        k0 = K["f00_Re" + t]
        f = work["rdata"][:, k0:k0 + nh]
        L = work["rdata"][:, K["L" + t]]
        mags = np.array([oracle.su_mag(nf, row) for row in f])
        before = np.array([oracle.su_mag(nf, row) for row in parts["rdata"][:, k0:k0 + nh]])
        fixed = np.abs(before - L) > 1e-5  # well above maxError: rescaled to |f| = L exactly
        assert fixed.sum() > 5
        np.testing.assert_allclose(mags[fixed], L[fixed], rtol=1e-13)
        assert np.abs(mags - L).max() <= 1e-6 * (1 + 1e-6)  # the rest were within maxError
        # the length rescale divides every component, so the trace ends near 1, not at 1
        assert np.abs(f[:, diag].sum(1) - 1).max() < 2e-4
    # a trace error above 100*maxError stops with OR_ERR_TRACE at the lowest index
    bad_parts = parts.copy()
    bad_parts["rdata"][3, K["f00_Re"]] += 1e-3
    bad_parts["rdata"][5, K["f00_Re"]] += 1e-3
    err, bad = oracle.renormalize(nf, p, bad_parts)
    assert err == oracle.ERR_TRACE and bad == 3
    # a negative diagonal between -100*maxError and -maxError is clipped to zero
    clip = parts.copy()
    k11 = K["f00_Re"] + herm_index(nf)["11_Re"]
    clip["rdata"][0, k11] = -5e-6
    clip["rdata"][0, K["f00_Re"]] += 5e-6
    oracle.renormalize(nf, p, clip)
    assert clip["rdata"][0, k11] == 0.0 or abs(clip["rdata"][0, k11]) < 1e-12


def test_redistribute_wraps_into_roundoff_domain(oracle):
    d = oracle.Deck(nf=2, ncell=(4, 4, 4), Lx=2.0, Ly=2.0, Lz=2.0)
    p = d.params()
    rhi = oracle.roundoff_hi(p)
    assert rhi[0] < 2.0 and math.floor(rhi[0] * 2.0) == 3
    parts = np.zeros(4, dtype=oracle.particle_dtype(2))
    parts["pos"] = [[2.0, -1e-17, 0.3], [2.1, -0.2, 1.9999999999], [0.0, 1.0, 2.0], [-2e-16, 0.5, 0.5]]
    oracle.redistribute(2, p, parts)
    pos = parts["pos"]
    assert pos[0, 0] == 0.0 and pos[0, 1] == rhi[1]
    assert pos[1, 0] == pytest.approx(0.1) and pos[1, 1] == pytest.approx(1.8)
    assert pos[2, 2] == 0.0 and pos[3, 0] == rhi[0]
    assert ((pos >= 0) & (pos <= 2.0)).all()


# --------------------------------------------------------------------------
# RK4: order of accuracy on the MSW problem (constant Hamiltonian)
# --------------------------------------------------------------------------
def test_rk4_fourth_order(oracle):
    deck = kat.reference_deck("inputs_msw_test")
    d = oracle.Deck(nf=2, simulation_type=0, ncell=(1, 1, 1), nphi_equator=1, rho_g_ccm=deck["rho_g_ccm"],
                    Ye=1.0, mass1_eV=deck["mass1_eV"], mass2_eV=0.0, theta12_degrees=deck["theta12_degrees"],
                    maxError=1.0)
    K = {nm: i for i, nm in enumerate(oracle.attribute_names(2))}
    T = 2.0e-11
    errs = []
    for nsteps in (20, 40):
        parts = oracle.init_particles(d)
        p = d.params()
        g = oracle.Grid(2, p)
        for _ in range(nsteps):
            oracle.rk4_step(2, p, parts, g, T / nsteps, 0.0)
        r = parts["rdata"][0]
        e = kat.msw_errors([r[K["time"]]], [r[K["f00_Re"]]], [r[K["f11_Re"]]], [r[K["f00_Rebar"]]],
                           [r[K["f11_Rebar"]]])
        errs.append(max(e["ee"], e["eebar"]))
    assert errs[0] / errs[1] == pytest.approx(16.0, rel=0.15)


@pytest.mark.parametrize("rk_type,order,n0", [(1, 1, 400), (2, 2, 40), (3, 3, 20), (4, 4, 20)])
def test_rk_tableau_orders(oracle, rk_type, order, n0):
    """Each preset tableau of integration.rk.type (the reference's convergence
    study, Scripts/tests/convergence.sh) converges at its order on MSW."""
    deck = kat.reference_deck("inputs_msw_test")
    d = oracle.Deck(nf=2, simulation_type=0, ncell=(1, 1, 1), nphi_equator=1, rho_g_ccm=deck["rho_g_ccm"],
                    Ye=1.0, mass1_eV=deck["mass1_eV"], mass2_eV=0.0, theta12_degrees=deck["theta12_degrees"],
                    maxError=1.0)
    K = {nm: i for i, nm in enumerate(oracle.attribute_names(2))}
    T = 2.0e-11
    errs = []
    for nsteps in (n0, 2 * n0):
        parts = oracle.init_particles(d)
        p = d.params()
        g = oracle.Grid(2, p)
        for _ in range(nsteps):
            oracle.rk_step(2, p, parts, g, T / nsteps, 0.0, rk_type)
        r = parts["rdata"][0]
        e = kat.msw_errors([r[K["time"]]], [r[K["f00_Re"]]], [r[K["f11_Re"]]], [r[K["f00_Rebar"]]],
                           [r[K["f11_Rebar"]]])
        errs.append(max(e["ee"], e["eebar"]))
    assert errs[0] / errs[1] == pytest.approx(2.0 ** order, rel=0.2)


def test_rk_step_rk4_tableau_is_rk4_step(oracle):
    d = cases.oracle_deck(oracle, cases.fast_flavor_dict(nf=3, ncell=(2, 2, 2), nphi=4, st4_amplitude=1e-3))
    a = oracle.init_particles(d)
    b = a.copy()
    p = d.params()
    ga, gb = oracle.Grid(3, p), oracle.Grid(3, p)
    ra = oracle.rk4_step(3, p, a, ga, 1e-12, 0.0)
    rb = oracle.rk_step(3, p, b, gb, 1e-12, 0.0, 4)
    assert ra == rb
    np.testing.assert_array_equal(a["rdata"], b["rdata"])


# --------------------------------------------------------------------------
# The reference's analytic acceptance tests (Scripts/tests, .travis.yml:32-36)
# --------------------------------------------------------------------------
def _run_kat(oracle, deckname, nsteps, nf=2, every=lambda p, K: None):
    deck = kat.reference_deck(deckname)
    kw = {k: deck[k] for k in oracle.Deck.__dataclass_fields__ if k in deck}
    kw["ncell"] = tuple(deck["ncell"])
    kw["nppc"] = tuple(deck["nppc"])
    d = oracle.Deck(nf=nf, **kw)
    parts = oracle.init_particles(d)
    p = d.params()
    g = oracle.Grid(nf, p)
    K = {nm: i for i, nm in enumerate(oracle.attribute_names(nf))}
    vv = oracle.vvac_max(nf, p, parts)
    oracle.deposit_to_mesh(nf, p, parts, g)
    err, dt = oracle.compute_dt(nf, p, g, vv)
    assert err == 0
    snaps = [every(parts, K)]
    for _ in range(nsteps):
        err, dt, bad = oracle.rk4_step(nf, p, parts, g, dt, vv)
        assert err == 0
        snaps.append(every(parts, K))
    return snaps


def _f01(parts, K, reduce=None):
    r = parts["rdata"]
    keys = ("time", "f01_Re", "f01_Im", "f01_Rebar", "f01_Imbar")
    if reduce is None:
        return [r[0, K[k]] for k in keys]
    return [r[0, K["time"]]] + [np.abs(r[:, K[k]]).max() for k in keys[1:]]


def test_kat_fast_flavor(oracle):
    s = np.array(_run_kat(oracle, "inputs_fast_flavor", 100, every=_f01))
    errs = kat.fast_flavor_errors(*s.T)
    assert max(errs) < 2e-2, errs


def test_kat_fast_flavor_nonzerok(oracle):
    s = np.array(_run_kat(oracle, "inputs_fast_flavor_nonzerok", 160, every=lambda p, K: _f01(p, K, "max")))
    errs = kat.fast_flavor_k_errors(*s.T)
    assert max(errs) < 1e-2, errs


def test_kat_msw(oracle):
    keys = ("time", "f00_Re", "f11_Re", "f00_Rebar", "f11_Rebar")
    s = np.array(_run_kat(oracle, "inputs_msw_test", 50, every=lambda p, K: [p["rdata"][0, K[k]] for k in keys]))
    e = kat.msw_errors(*s.T)
    assert max(e.values()) < 7e-2, e


def test_kat_bipolar_runs(oracle):
    _run_kat(oracle, "inputs_bipolar_test", 40, every=lambda p, K: None)


# ---------------------------------------------------------------------------
# initial-condition helpers
# ---------------------------------------------------------------------------
def test_philox_known_answers(oracle):
    """Philox4x32-10 against the known-answer vectors published with Random123
    (Salmon et al., SC'11; kat_vectors: counter, key -> output)."""
    kats = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
            ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
            ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
             (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]
    for ctr, key, want in kats:
        got = oracle.philox4x32(*ctr, *key)
        assert [int(x) for x in got] == list(want)


def test_uniform01_range_and_moments(oracle):
    u = oracle.uniform01(99, np.arange(200000, dtype=np.uint64), 3)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 3e-3 and abs(u.var() - 1.0 / 12.0) < 2e-3
    # the two halves of one Philox block are different draws
    v = oracle.uniform01(99, np.arange(1000, dtype=np.uint64), 2)
    w = oracle.uniform01(99, np.arange(1000, dtype=np.uint64), 3)
    assert not np.any(v == w)


def test_minerbo_closure_solves_its_equation(oracle):
    """Minerbo: the flux factor of the closure is coth(Z) - 1/Z
    (FlavoredNeutrinoContainerInit.cpp:46-80); below 1e-3 the reference uses
    the small-Z limit Z = 3 ff."""
    for ff in (0.01, 0.1, 0.33, 0.6, 0.9):
        Z = oracle.minerbo_Z(ff)
        assert abs(1.0 / math.tanh(Z) - 1.0 / Z - ff) < 1e-6
    assert oracle.minerbo_Z(5e-4) == 3 * 5e-4
    assert oracle.minerbo_Z(0.0) == 0.0


def test_uniform_sphere_counts_and_antipodes(oracle):
    """uniform_sphere_xyz(16) is the 92-direction set of inputs_1d_fiducial;
    every direction off the equator comes with its antipode and all are unit
    vectors."""
    for nphi, n in ((16, 92), (32, 378)):
        xyz = oracle.uniform_sphere_xyz(nphi)
        assert len(xyz) == n
        np.testing.assert_allclose(np.linalg.norm(xyz, axis=1), 1.0, atol=1e-15)
        off = xyz[np.abs(xyz[:, 2]) > 1e-12]
        for v in off[:20]:
            assert np.min(np.linalg.norm(off + v, axis=1)) < 1e-15


def test_sync_location_and_apply_rhs(oracle):
    """SyncLocation copies between pos and (x,y,z) (FlavoredNeutrinoContainer.cpp:8-41);
    ApplyFlavoredNeutrinoRHS is p = a dF + p on every attribute (.H:33-47)."""
    d = cases.oracle_deck(oracle, cases.fast_flavor_dict(nf=2, ncell=(2, 2, 2), nphi=4))
    parts = oracle.init_particles(d)
    K = {n: i for i, n in enumerate(oracle.attribute_names(2))}
    # type 0 = CoordinateToPosition: pos <- (x, y, z)
    p = parts.copy()
    p["rdata"][:, K["x"]] = -7.0
    oracle.sync_location(2, p, 0)
    np.testing.assert_array_equal(p["pos"][:, 0], -7.0)
    np.testing.assert_array_equal(p["pos"][:, 1:], parts["pos"][:, 1:])
    # type 1 = PositionToCoordinate: (x, y, z) <- pos
    q = parts.copy()
    q["pos"] += 0.125
    oracle.sync_location(2, q, 1)
    np.testing.assert_array_equal(q["rdata"][:, K["x"]:K["z"] + 1], parts["pos"] + 0.125)
    r = parts.copy()
    dF = parts.copy()
    dF["rdata"] = np.random.default_rng(1).standard_normal(dF["rdata"].shape)
    oracle.apply_rhs(2, r, 0.25, dF)
    np.testing.assert_array_equal(r["rdata"], 0.25 * dF["rdata"] + parts["rdata"])


# --------------------------------------------------------------------------
# initial conditions: angular moments of the generated distributions
# --------------------------------------------------------------------------
def deposited_moments(oracle, deck, parts, bar, comp):
    """(total N_comp, total F_comp vector) over the domain's cells, per unit
    volume, from the oracle's deposit."""
    nf = deck.nf
    p = deck.params()
    g = oracle.Grid(nf, p)
    oracle.deposit_to_mesh(nf, p, parts, g)
    v = g.valid()
    nh = nf * nf
    vol = deck.Lx * deck.Ly * deck.Lz
    tot = [v[3 + (2 * var + bar) * nh + comp].sum() / vol for var in range(4)]
    return tot[0], np.array(tot[1:])


@pytest.mark.parametrize("directions,tol", [("cube26", 1e-13), ("sphere", 2e-3)])
def test_init_type4_moments(oracle, directions, tol):
    """Sim type 4 weights each direction by n dV/ndirs (1 + 3 ff cos theta)
    (FlavoredNeutrinoContainerInit.cpp:506-569).  For a direction set whose
    second moment is isotropic, <u u> = 1/3, the deposit's number density is
    n and its flux F = n ff along the flux axis: |F|/N is the flux factor.
    The 26 cube directions satisfy <u u> = 1/3 exactly; the 378 of
    uniform_sphere_xyz(32) to 1e-3."""
    d = cases.fast_flavor_dict(nf=2, ncell=(2, 2, 2), nphi=32, directions=directions, st4_fluxfac=0.2,
                               st4_theta=0.7, st4_phi=1.9, st4_fluxfacbar=0.33, st4_amplitude=0.0)
    deck = cases.oracle_deck(oracle, d)
    parts = oracle.init_particles(deck)
    for bar, n, ff, th, ph in ((0, 4.89e32, 0.2, 0.7, 1.9), (1, 4.89e32, 0.33, 0.0, 0.0)):
        N, F = deposited_moments(oracle, deck, parts, bar, 0)
        axis = np.array([math.cos(ph) * math.sin(th), math.sin(ph) * math.sin(th), math.cos(th)])
        assert abs(N / n - 1) < 1e-13
        np.testing.assert_allclose(F / N, ff * axis, atol=tol * ff)


@pytest.mark.parametrize("nf", [2, 3])
def test_init_type5_minerbo_moments(oracle, nf):
    """Sim type 5 weights each direction with the Minerbo closure
    exp(Z mu) Z/sinh(Z) (FlavoredNeutrinoContainerInit.cpp:574-654), whose
    integral over the sphere is 1 and whose mean mu is coth(Z) - 1/Z = ff.
    So the deposit of nu_e N_ee is n_nue with flux n_nue ff_e, the same for
    nubar_e, and each heavy-lepton diagonal holds n_nux/4 with its flux --
    to the 378-direction quadrature's accuracy (1e-3)."""
    d = cases.minerbo_dict(nf=nf, st5_amplitude=0.0)
    deck = cases.oracle_deck(oracle, d)
    parts = oracle.init_particles(deck)
    herm = herm_index(nf)
    flux = {"nue": np.array([d["st5_fxnue"], d["st5_fynue"], d["st5_fznue"]]),
            "nua": np.array([d["st5_fxnua"], d["st5_fynua"], d["st5_fznua"]]),
            "nux": np.array([d["st5_fxnux"], d["st5_fynux"], d["st5_fznux"]])}
    checks = [(0, "00_Re", "nue", 1.0), (1, "00_Re", "nua", 1.0), (0, "11_Re", "nux", 0.25), (1, "11_Re", "nux", 0.25)]
    if nf == 3:
        checks += [(0, "22_Re", "nux", 0.25), (1, "22_Re", "nux", 0.25)]
    for bar, comp, fl, frac in checks:
        N, F = deposited_moments(oracle, deck, parts, bar, herm[comp])
        n = d["st5_n" + fl] * frac
        assert abs(N / n - 1) < 2e-3, (bar, comp, N / n)
        np.testing.assert_allclose(F / N, flux[fl], atol=2e-3 * max(np.linalg.norm(flux[fl]), 1e-3))
    # each particle's flavour state is a density matrix: trace 1
    K = {nm: i for i, nm in enumerate(oracle.attribute_names(nf))}
    for t in ("", "bar"):
        tr = sum(parts["rdata"][:, K[f"f{i}{i}_Re{t}"]] for i in range(nf))
        np.testing.assert_allclose(tr, 1.0, atol=1e-15)
