"""Parity of the HIP paths with the FP64 oracle on identical inputs.

Bars: cell indices and positions bit-exact where the arithmetic allows it;
FP64 fields within 1e-12 relative to each component's max-abs (per-element
relative error is ill-defined for cancelled fluxes); FP32-bulk within the
tolerance stated in each test.
"""

import numpy as np
import pytest

import cases
import kat

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def core():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from emu_amd import core as c
    return c


def random_case(oracle, nf, ncell, L, n, seed):
    """Random positions, directions, energies, weights and flavour states."""
    rng = np.random.default_rng(seed)
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
        k0 = K["f00_Re" + t]
        rd[:, k0:k0 + nh] = rng.normal(size=(n, nh)) * 0.1
    parts["id"] = np.arange(1, n + 1)
    return parts, K


def make(core, oracle, d):
    deck = cases.oracle_deck(oracle, d)
    emu = core.Emu(cases.product_params(d))
    return deck, deck.params(), emu


# ---------------------------------------------------------------------------
# per-function entry points
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("nf", [2, 3])
def test_init_matches_oracle(core, oracle, nf):
    d = cases.fast_flavor_dict(nf=nf, ncell=(3, 4, 5), nphi=8, st3_amplitude=1e-5)
    deck, prm, emu = make(core, oracle, d)
    want = oracle.init_particles(deck)
    got = core.FlavoredNeutrinoContainer.init_particles(emu).to_aos()
    assert len(got) == len(want)
    np.testing.assert_array_equal(got["id"], want["id"])
    np.testing.assert_array_equal(got["pos"], want["pos"])
    assert cases.max_rel_err(got["rdata"], want["rdata"]) < 1e-14


def sim_type_dict(st, nf):
    """A deck for each initial condition: the reference's own sample decks for
    types 0-3 (MSW, bipolar, fast flavor, fast flavor with k != 0), a
    fast-flavor box with a tilted flux and a single mode for 4, and the
    Minerbo closure for 5."""
    if st == 4:
        return cases.fast_flavor_dict(nf=nf, ncell=(3, 2, 4), nphi=8, st4_fluxfac=0.2, st4_theta=0.7, st4_phi=1.9,
                                      st3_amplitude=1e-5)
    if st == 5:
        return cases.minerbo_dict(nf=nf, ncell=(2, 3, 2), nphi=16)
    name = {0: "inputs_msw_test", 1: "inputs_bipolar_test", 2: "inputs_fast_flavor",
            3: "inputs_fast_flavor_nonzerok"}[st]
    d = cases.deck_dict(name, nf)
    assert d["simulation_type"] == st
    return d


@pytest.mark.parametrize("st", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("nf", [2, 3])
def test_init_sim_types_match_oracle(core, oracle, nf, st):
    """InitParticles (FlavoredNeutrinoContainerInit.cpp:136-661) for every
    simulation type: the GPU generator equals the oracle's bit for bit in
    ids and positions, and to 1e-14 in the attributes (the generator's
    transcendental functions are the device's, not glibc's)."""
    d = sim_type_dict(st, nf)
    deck, prm, emu = make(core, oracle, d)
    want = oracle.init_particles(deck)
    got = core.FlavoredNeutrinoContainer.init_particles(emu).to_aos()
    assert len(got) == len(want) > 0
    np.testing.assert_array_equal(got["id"], want["id"])
    np.testing.assert_array_equal(got["pos"], want["pos"])
    assert cases.max_rel_err(got["rdata"], want["rdata"]) < 1e-14


@pytest.mark.parametrize("nf", [2, 3])
# (13, 2, 3): fab rows longer than the gather's 8-box runs, the last run partial
@pytest.mark.parametrize("ncell", [(4, 3, 5), (1, 1, 6), (2, 2, 2), (13, 2, 3)])
def test_deposit_fill_interpolate_parity(core, oracle, nf, ncell):
    L = tuple(0.5 * n for n in ncell)
    d = cases.fast_flavor_dict(nf=nf, ncell=ncell, L=L, rho_g_ccm=1e10, Ye=0.45)
    deck, prm, emu = make(core, oracle, d)
    parts, K = random_case(oracle, nf, ncell, L, 300, seed=nf * 100 + ncell[2])
    g_or = oracle.Grid(nf, prm)
    oracle.deposit_to_mesh(nf, prm, parts, g_or)
    c = core.FlavoredNeutrinoContainer.from_aos(emu, parts)
    state = core.MultiFab(emu)
    core.deposit_to_mesh(c, state)
    fab = state.data.cpu().numpy().reshape(emu.ncomp, -1).T
    want = g_or.fab().reshape(emu.ncomp, -1).T
    # deposit components over valid and ghost cells (ghosts keep SumBoundary partials)
    assert cases.max_rel_err(fab[:, 3:], want[:, 3:]) < 1e-12
    oracle.fill_boundary(nf, prm, g_or)
    state.FillBoundary()
    fab = state.data.cpu().numpy().reshape(emu.ncomp, -1).T
    want = g_or.fab().reshape(emu.ncomp, -1).T
    assert cases.max_rel_err(fab, want) < 1e-12
    # interpolate_rhs_from_mesh on the oracle's own (filled) grid
    state.data.copy_(__import__("torch").from_numpy(g_or.fab().copy()))
    rhs_or = parts.copy()
    oracle.interpolate_rhs_from_mesh(nf, prm, rhs_or, g_or)
    rhs = core.FlavoredNeutrinoContainer.from_aos(emu, parts)
    core.interpolate_rhs_from_mesh(rhs, state)
    got = rhs.to_aos()["rdata"]
    assert cases.max_rel_err(got, rhs_or["rdata"]) < 1e-12
    # compute_dt over the fab box
    vv = oracle.vvac_max(nf, prm, parts)
    assert c.Vvac_max() == pytest.approx(vv, rel=1e-15)
    err, dt_or = oracle.compute_dt(nf, prm, g_or, vv)
    assert err == 0
    assert core.compute_dt(state, vv) == pytest.approx(dt_or, rel=1e-14)


@pytest.mark.parametrize("nf", [2, 3])
def test_deposit_interpolate_ghost_shell_and_skipped(core, oracle, nf):
    """Particles between SyncLocation and RedistributeLocal may sit outside the
    domain.  Those within one cell of it (a third of them here, in random
    order) deposit into the fab's ghost cells, which SumBoundary folds, and
    gather from them, as in the oracle.  Two particles three cells out have
    stencils that leave the fab: both calls skip them, and emu_sync reports
    them (the reference would write outside its fab)."""
    ncell, L = (4, 3, 5), (2.0, 1.5, 2.5)
    d = cases.fast_flavor_dict(nf=nf, ncell=ncell, L=L, rho_g_ccm=1e10, Ye=0.45)
    deck, prm, emu = make(core, oracle, d)
    parts, K = random_case(oracle, nf, ncell, L, 400, seed=17 + nf)
    rng = np.random.default_rng(5)
    dx = np.array(L) / np.array(ncell)
    shell = rng.choice(len(parts), size=len(parts) // 3, replace=False)
    for i in shell:
        dim = rng.integers(3)
        side = rng.choice([-1.0, 1.0])
        x = -rng.uniform(0, 1) * dx[dim] if side < 0 else L[dim] + rng.uniform(0, 0.999) * dx[dim]
        parts["pos"][i, dim] = x
        parts["rdata"][i, 1 + dim] = x
    far = parts[:2].copy()
    far["pos"][:, 0] = -3.2 * dx[0]
    far["rdata"][:, 1] = -3.2 * dx[0]
    g_or = oracle.Grid(nf, prm)
    oracle.deposit_to_mesh(nf, prm, parts, g_or)
    c = core.FlavoredNeutrinoContainer.from_aos(emu, np.concatenate([parts, far]))
    state = core.MultiFab(emu)
    core.deposit_to_mesh(c, state)
    with pytest.raises(Exception) as ei:
        emu.sync()
    assert ei.value.code == 10 and "2 particle" in str(ei.value)
    fab = state.data.cpu().numpy().reshape(emu.ncomp, -1).T
    want = g_or.fab().reshape(emu.ncomp, -1).T
    assert cases.max_rel_err(fab[:, 3:], want[:, 3:]) < 1e-12
    oracle.fill_boundary(nf, prm, g_or)
    state.data.copy_(__import__("torch").from_numpy(g_or.fab().copy()))
    rhs_or = parts.copy()
    oracle.interpolate_rhs_from_mesh(nf, prm, rhs_or, g_or)
    rhs = core.FlavoredNeutrinoContainer.from_aos(emu, np.concatenate([parts, far]))
    core.interpolate_rhs_from_mesh(rhs, state)
    with pytest.raises(Exception):
        emu.sync()
    got = rhs.to_aos()
    assert cases.max_rel_err(got["rdata"][:len(parts)], rhs_or["rdata"]) < 1e-12
    # the skipped particles keep their attributes (the rhs container's copy)
    np.testing.assert_array_equal(got["rdata"][len(parts):], far["rdata"])


@pytest.mark.parametrize("nf", [2, 3])
def test_renormalize_sync_wrap_apply_parity(core, oracle, nf):
    d = cases.fast_flavor_dict(nf=nf, ncell=(4, 4, 4), nphi=4, st4_amplitude=1e-3)
    deck, prm, emu = make(core, oracle, d)
    parts = oracle.init_particles(deck)
    K = {nm: i for i, nm in enumerate(oracle.attribute_names(nf))}
    nh = nf * nf
    rng = np.random.default_rng(4)
    for t in ("", "bar"):
        k0 = K["f00_Re" + t]
        parts["rdata"][:, k0:k0 + nh] *= 1 + rng.uniform(-3e-5, 3e-5, size=(len(parts), 1))
    # positions pushed across the boundary for the wrap
    parts["rdata"][:, K["x"]] += rng.choice([-2.5, 0.0, 2.5], size=len(parts))
    c = core.FlavoredNeutrinoContainer.from_aos(emu, parts)
    oracle.renormalize(nf, prm, parts)
    c.Renormalize()
    for tp in (0,):
        oracle.sync_location(nf, parts, tp)
        c.SyncLocation(tp)
    oracle.redistribute(nf, prm, parts)
    c.RedistributeLocal()
    oracle.sync_location(nf, parts, 1)
    c.SyncLocation(1)
    got = c.to_aos()
    np.testing.assert_array_equal(got["pos"], parts["pos"])
    assert cases.max_rel_err(got["rdata"], parts["rdata"]) < 1e-15
    # apply_rhs is a plain saxpy: bit-exact
    dF = core.FlavoredNeutrinoContainer.from_aos(emu, parts)
    oracle.apply_rhs(nf, parts, 1.7e-12, parts.copy())
    c.apply_rhs(1.7e-12, dF)
    np.testing.assert_array_equal(c.to_aos()["rdata"], parts["rdata"])
    # error path: a trace error beyond 100*maxError reports the lowest index
    bad = parts.copy()
    bad["rdata"][7, K["f00_Re"]] += 1e-3
    bad["rdata"][9, K["f00_Re"]] += 1e-3
    cb = core.FlavoredNeutrinoContainer.from_aos(emu, bad)
    with pytest.raises(Exception) as ei:
        cb.Renormalize()
    assert ei.value.code == 1 and ei.value.bad_index == 7


# ---------------------------------------------------------------------------
# whole steps
# ---------------------------------------------------------------------------
def oracle_steps(oracle, nf, prm, parts, nsteps, fixed_dt=None):
    g = oracle.Grid(nf, prm)
    vv = oracle.vvac_max(nf, prm, parts)
    oracle.deposit_to_mesh(nf, prm, parts, g)
    err, dt = oracle.compute_dt(nf, prm, g, vv)
    assert err == 0
    dts = []
    for _ in range(nsteps):
        h = fixed_dt if fixed_dt else dt
        dts.append(h)
        err, dt, bad = oracle.rk4_step(nf, prm, parts, g, h, vv)
        assert err == 0
    return dts, dt


@pytest.mark.parametrize("n", [0, 1, 3, 4])
def test_rk_combine_is_copy_then_saxpys(core, oracle, n):
    """emu_rk_combine equals copyParticles followed by one apply_rhs per stage,
    bit for bit (the same roundings in the same order)."""
    import torch
    nf = 3
    d = cases.fast_flavor_dict(nf=nf, ncell=(3, 3, 3), nphi=4)
    deck, prm, emu = make(core, oracle, d)
    parts, K = random_case(oracle, nf, (3, 3, 3), (1.5, 1.5, 1.5), 500, seed=21)
    src = core.FlavoredNeutrinoContainer.from_aos(emu, parts)
    gen = torch.Generator(device="cpu").manual_seed(n)
    Fs = []
    for _ in range(n):
        F = core.FlavoredNeutrinoContainer(emu, src.np)
        F.rdata.copy_(torch.randn(F.rdata.shape, generator=gen, dtype=torch.float64))
        Fs.append(F)
    a = [1.3e-12 * (j + 1) for j in range(n)]
    want = core.FlavoredNeutrinoContainer(emu, src.np)
    want.copyParticles(src)
    for aj, F in zip(a, Fs):
        want.apply_rhs(aj, F)
    got = core.FlavoredNeutrinoContainer(emu, src.np)
    got.rk_combine(src, a, Fs)
    assert torch.equal(got.rdata, want.rdata) and torch.equal(got.pos, want.pos)
    assert torch.equal(got.ids, src.ids)


@pytest.mark.parametrize("nf", [2, 3])
def test_integrator_parity(core, oracle, nf):
    d = cases.fast_flavor_dict(nf=nf, ncell=(4, 4, 4), nphi=4)
    deck, prm, emu = make(core, oracle, d)
    parts = oracle.init_particles(deck)
    c = core.FlavoredNeutrinoContainer.from_aos(emu, parts)
    integ = core.Integrator(emu, c)
    dts, dt_next = oracle_steps(oracle, nf, prm, parts, 4)
    out = integ.integrate(4).to_aos()
    assert integ.dt == pytest.approx(dt_next, rel=1e-13)
    np.testing.assert_allclose(out["pos"], parts["pos"], rtol=1e-14, atol=0)
    assert cases.max_rel_err(out["rdata"], parts["rdata"]) < 1e-12


@pytest.mark.parametrize("nf", [2, 3])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_flavor_amplitude_diagnostic(core, oracle, nf, prec, tol):
    """emu_state_flavor_amplitude: sums over cells of |N_01| and |Nbar_01| of
    the deposit (the growth-rate diagnostic) against the oracle's fab."""
    d = cases.fast_flavor_dict(nf=nf, ncell=(4, 4, 4), nphi=4, st4_amplitude=1e-3)
    deck, prm, emu = make(core, oracle, d)
    parts = oracle.init_particles(deck)
    c = core.FlavoredNeutrinoContainer.from_aos(emu, parts)
    st = core.FusedState(emu, c.np, prec)
    st.load(c)
    got = st.flavor_amplitude()
    st.close()
    g = oracle.Grid(nf, prm)
    oracle.deposit_to_mesh(nf, prm, parts, g)
    K = {n: i for i, n in enumerate(oracle.grid_names(nf))}
    v = g.valid()
    a = np.sqrt(v[K["N01_Re"]] ** 2 + v[K["N01_Im"]] ** 2).sum()
    b = np.sqrt(v[K["N01_Rebar"]] ** 2 + v[K["N01_Imbar"]] ** 2).sum()
    assert got[0] == pytest.approx(a, rel=tol) and got[1] == pytest.approx(b, rel=tol)


@pytest.mark.parametrize("itype,rk", [(0, 1), (1, 2), (1, 3)])
def test_integrator_tableaus_parity(core, oracle, itype, rk):
    """Forward Euler (integration.type = 0), trapezoid and SSPRK3 on the
    per-function path against the oracle's tableau step, cfl_factor = -1 as
    in the reference's convergence study."""
    nf = 3
    d = cases.fast_flavor_dict(nf=nf, ncell=(4, 4, 4), nphi=4, cfl_factor=-1.0)
    d["integration.type"], d["integration.rk.type"] = itype, rk
    deck, prm, emu = make(core, oracle, d)
    parts = oracle.init_particles(deck)
    c = core.FlavoredNeutrinoContainer.from_aos(emu, parts)
    integ = core.Integrator(emu, c)
    g = oracle.Grid(nf, prm)
    vv = oracle.vvac_max(nf, prm, parts)
    oracle.deposit_to_mesh(nf, prm, parts, g)
    err, dt = oracle.compute_dt(nf, prm, g, vv)
    assert err == 0 and integ.dt == pytest.approx(dt, rel=1e-13)
    for _ in range(3):
        err, dt, bad = oracle.rk_step(nf, prm, parts, g, dt, vv, rk)
        assert err == 0
    out = integ.integrate(3).to_aos()
    assert integ.dt == pytest.approx(dt, rel=1e-13)
    np.testing.assert_allclose(out["pos"], parts["pos"], rtol=1e-14, atol=0)
    assert cases.max_rel_err(out["rdata"], parts["rdata"]) < 1e-12
    with pytest.raises(ValueError):
        core.FusedState(emu, c.np, "fp64")


@pytest.mark.parametrize("nf", [2, 3])
@pytest.mark.parametrize("ncell", [(4, 4, 4), (1, 1, 8), (3, 5, 2), (2, 3, 7), (1, 1, 5)])
def test_fused_fp64_parity(core, oracle, nf, ncell):
    """Cells are taken four to a workgroup (2 x 2 in x, y, or 4 along z): odd
    counts leave partial patches, 2-cell dimensions make a patch's
    neighbourhood hold periodic images twice."""
    d = cases.fast_flavor_dict(nf=nf, ncell=ncell, nphi=8)
    deck, prm, emu = make(core, oracle, d)
    parts = oracle.init_particles(deck)
    # a fixed dt large enough that particles change cells during the run
    fixed = 0.3 * 0.5 / oracle.clight
    c = core.FlavoredNeutrinoContainer.from_aos(emu, parts)
    st = core.FusedState(emu, c.np, "fp64")
    st.load(c)
    st.evolve(6, fixed_dt=fixed)
    st.store(c)
    info = st.info()
    oracle_steps(oracle, nf, prm, parts, 6, fixed_dt=fixed)
    got = cases.by_id(c.to_aos())
    want = cases.by_id(parts)
    np.testing.assert_array_equal(got["pos"], want["pos"])  # positions and cell binning bit-exact
    assert cases.max_rel_err(got["rdata"], want["rdata"]) < 1e-12
    assert info["steps"] == 6
    assert info["time"] == pytest.approx(want["rdata"][0, 0], rel=1e-15)


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("nf", [2, 3])
def test_lower_shape_orders_parity(core, oracle, nf, order):
    """SHAPE_FACTOR_ORDER 0 (nearest cell: 1-point stencils) and 1 (cloud in
    cell: the 3-point stencils with one zero weight) on the per-function
    deposit and interpolation and on 6 fused FP64 steps with a dt that moves
    particles across cells, against the oracle at the same order.  The
    densities are a tenth of the fast-flavor deck's, so that this dt stays
    within compute_dt's flavour bound: at the full densities it is 5x that
    bound, and the oracle's own step then grows a one-ulp change of the
    flavour state to ~3e-11 in 6 steps at these orders (1e-15 here), which
    no 1e-12 comparison could survive."""
    d = cases.fast_flavor_dict(nf=nf, ncell=(4, 3, 5), L=(2.0, 1.5, 2.5), nphi=8, shape_order=order,
                               st4_amplitude=1e-3, st4_ndens=4.89e31, st4_ndensbar=4.89e31)
    deck, prm, emu = make(core, oracle, d)
    parts = oracle.init_particles(deck)
    c = core.FlavoredNeutrinoContainer.from_aos(emu, parts)
    g_or = oracle.Grid(nf, prm)
    oracle.deposit_to_mesh(nf, prm, parts, g_or)
    oracle.fill_boundary(nf, prm, g_or)
    state = core.MultiFab(emu)
    core.deposit_to_mesh(c, state)
    state.FillBoundary()
    fab = state.data.cpu().numpy().reshape(emu.ncomp, -1)
    want = g_or.fab().reshape(emu.ncomp, -1)
    # the isotropic direction set's net fluxes cancel to rounding (~1e-16 of N), so
    # N and F are measured against the largest of them (|F| <= N)
    np.testing.assert_array_equal(fab[:3], want[:3])
    assert np.abs(fab[3:] - want[3:]).max() / np.abs(want[3:]).max() < 1e-12
    state.data.copy_(__import__("torch").from_numpy(g_or.fab().copy()))
    rhs_or = parts.copy()
    oracle.interpolate_rhs_from_mesh(nf, prm, rhs_or, g_or)
    rhs = core.FlavoredNeutrinoContainer.from_aos(emu, parts)
    core.interpolate_rhs_from_mesh(rhs, state)
    assert cases.max_rel_err(rhs.to_aos()["rdata"], rhs_or["rdata"]) < 1e-12
    fixed = 0.3 * 0.5 / oracle.clight
    st = core.FusedState(emu, c.np, "fp64")
    st.load(c)
    st.evolve(6, fixed_dt=fixed)
    st.store(c)
    st.close()
    oracle_steps(oracle, nf, prm, parts, 6, fixed_dt=fixed)
    got = cases.by_id(c.to_aos())
    want = cases.by_id(parts)
    np.testing.assert_array_equal(got["pos"], want["pos"])
    assert cases.max_rel_err(got["rdata"], want["rdata"]) < 1e-12


@pytest.mark.parametrize("nf", [2, 3])
def test_fused_fp64_computed_dt(core, oracle, nf):
    d = cases.fast_flavor_dict(nf=nf, ncell=(4, 4, 4), nphi=4)
    deck, prm, emu = make(core, oracle, d)
    parts = oracle.init_particles(deck)
    c = core.FlavoredNeutrinoContainer.from_aos(emu, parts)
    st = core.FusedState(emu, c.np, "fp64")
    st.load(c)
    dt0 = st.info()["dt"]
    dts, dt_next = oracle_steps(oracle, nf, prm, parts, 5)
    assert dt0 == pytest.approx(dts[0], rel=1e-14)
    st.evolve(5)
    st.store(c)
    assert st.info()["dt"] == pytest.approx(dt_next, rel=1e-13)
    got = cases.by_id(c.to_aos())
    want = cases.by_id(parts)
    np.testing.assert_allclose(got["pos"], want["pos"], rtol=1e-14)
    assert cases.max_rel_err(got["rdata"], want["rdata"]) < 1e-11


@pytest.mark.parametrize("speedup", [4.0, 1e9])
@pytest.mark.parametrize("path", ["integrator", "fp64", "fp32"])
@pytest.mark.parametrize("nf", [2, 3])
def test_adaptive_dt_parity(core, oracle, nf, path, speedup):
    """max_adaptive_speedup > 1 (Evolve.cpp:68-74): dt = min(speedup dt_stupid,
    dt_adaptive).  2 cm cells keep the translation bound (cfl_factor 0.9, as
    the fused sort needs c dt < dx) above both, and the antineutrino flux makes
    |F - Fbar| the adaptive bound: a speedup of 4 is capped by dt_stupid, 1e9
    by dt_adaptive.  Every computed dt of 3 steps matches.  At 1e9 the step is
    far past the conservative bound, and the oracle's own 3 steps grow a
    one-ulp change of the flavour state to 1.2e-11 (NF=2) / 1.6e-11 (NF=3)
    of each component's range (9e-16 at 4): the FP64 state bar there is 1e-9,
    a hundred such ulps."""
    ncell = (4, 3, 5)
    d = cases.fast_flavor_dict(nf=nf, ncell=ncell, L=tuple(2.0 * n for n in ncell), nphi=4, st4_amplitude=1e-3,
                               cfl_factor=0.9, max_adaptive_speedup=speedup)
    deck, prm, emu = make(core, oracle, d)
    parts = oracle.init_particles(deck)
    c = core.FlavoredNeutrinoContainer.from_aos(emu, parts)
    probe = oracle.Deck(**{**deck.__dict__, "max_adaptive_speedup": 0.0}).params()
    g = oracle.Grid(nf, prm)
    oracle.deposit_to_mesh(nf, prm, parts, g)
    vv = oracle.vvac_max(nf, prm, parts)
    _, dt_stupid = oracle.compute_dt(nf, probe, g, vv)
    dts, dt_next = oracle_steps(oracle, nf, prm, parts, 3)
    if speedup == 4.0:
        assert dts[0] == pytest.approx(4.0 * dt_stupid, rel=1e-13)
    else:
        assert dts[0] > 4.0 * dt_stupid and dts[0] < 0.9 * 2.0 / oracle.clight
    ftol = 1e-13 if path != "fp32" else 1e-5
    if path == "integrator":
        integ = core.Integrator(emu, c)
        assert integ.dt == pytest.approx(dts[0], rel=1e-14)
        out = integ.integrate(3)
        assert integ.dt == pytest.approx(dt_next, rel=ftol)
        got = cases.by_id(out.to_aos())
    else:
        st = core.FusedState(emu, c.np, path)
        st.load(c)
        assert st.info()["dt"] == pytest.approx(dts[0], rel=1e-14 if path == "fp64" else 1e-6)
        st.evolve(3)
        st.store(c)
        assert st.info()["dt"] == pytest.approx(dt_next, rel=ftol)
        st.close()
        got = cases.by_id(c.to_aos())
    want = cases.by_id(parts)
    if path != "fp32":
        assert cases.max_rel_err(got["rdata"], want["rdata"]) < (1e-11 if speedup == 4.0 else 1e-9)
        return
    # At the adaptive step (31x the stupid one) the oracle itself moves 5-8% of the
    # largest |f11| (~1e-5) when its input is perturbed by 6e-8 relative, so
    # FP32-bulk is held to the FP32 bar against the trace Tr f = 1 of each
    # density matrix, and per attribute elsewhere.
    K = emu.K
    fcols = [K[n] for n in K if n.startswith("f")]
    other = [i for i in range(want["rdata"].shape[1]) if i not in fcols]
    assert cases.max_rel_err(got["rdata"][:, other], want["rdata"][:, other]) < 2e-5
    assert np.abs(got["rdata"][:, fcols] - want["rdata"][:, fcols]).max() < 2e-5


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("fp32", 2e-5)])
@pytest.mark.parametrize("nf", [2, 3])
def test_fused_full_chunks(core, oracle, nf, prec, tol):
    """92 particles per cell (uniform_sphere_xyz(16), the benchmark's direction
    set): every cell fills one 64-particle chunk and part of a second, so each
    row of the per-wave LDS staging is used; computed dt, so the auxiliary
    deposit runs too."""
    d = cases.fast_flavor_dict(nf=nf, ncell=(4, 4, 4), nphi=16, st4_amplitude=1e-3)
    deck, prm, emu = make(core, oracle, d)
    parts = oracle.init_particles(deck)
    assert len(parts) == 64 * 92
    c = core.FlavoredNeutrinoContainer.from_aos(emu, parts)
    st = core.FusedState(emu, c.np, prec)
    st.load(c)
    dt0 = st.info()["dt"]
    dts, dt_next = oracle_steps(oracle, nf, prm, parts, 4)
    assert dt0 == pytest.approx(dts[0], rel=1e-14 if prec == "fp64" else 1e-6)
    st.evolve(4)
    st.store(c)
    assert st.info()["dt"] == pytest.approx(dt_next, rel=1e-13 if prec == "fp64" else 1e-5)
    st.close()
    got = cases.by_id(c.to_aos())
    want = cases.by_id(parts)
    np.testing.assert_allclose(got["pos"], want["pos"], rtol=1e-14 if prec == "fp64" else 1e-7)
    assert cases.max_rel_err(got["rdata"], want["rdata"]) < tol


@pytest.mark.parametrize("nf", [2, 3])
def test_fused_fp32_bulk_tolerance(core, oracle, nf):
    """FP32 bulk (FP64 positions, deposit accumulation, Renormalize, dt): fields
    within 2e-5 of each component's max-abs after 5 steps."""
    d = cases.fast_flavor_dict(nf=nf, ncell=(4, 4, 4), nphi=4)
    deck, prm, emu = make(core, oracle, d)
    parts = oracle.init_particles(deck)
    c = core.FlavoredNeutrinoContainer.from_aos(emu, parts)
    st = core.FusedState(emu, c.np, "fp32")
    st.load(c)
    dts, _ = oracle_steps(oracle, nf, prm, parts, 5)
    st.evolve(5)
    st.store(c)
    got = cases.by_id(c.to_aos())
    want = cases.by_id(parts)
    np.testing.assert_allclose(got["pos"], want["pos"], rtol=1e-7)
    assert cases.max_rel_err(got["rdata"], want["rdata"]) < 2e-5


# ---------------------------------------------------------------------------
# the reference's analytic acceptance tests on every HIP path
# ---------------------------------------------------------------------------
def run_kat(core, deckname, nsteps, path, keys, reduce=None):
    d = cases.deck_dict(deckname)
    emu = core.Emu(cases.product_params(d))
    c = core.FlavoredNeutrinoContainer.init_particles(emu)
    K = emu.K
    snaps = []

    def snap(arr):
        r = arr["rdata"]
        if reduce is None:
            r0 = r[np.argmin(arr["id"])]
            return [r0[K[k]] for k in keys]
        return [r[0, K[keys[0]]]] + [np.abs(r[:, K[k]]).max() for k in keys[1:]]

    snaps.append(snap(c.to_aos()))
    if path == "integrator":
        integ = core.Integrator(emu, c)
        for _ in range(nsteps):
            integ.advance()
            snaps.append(snap(integ.old.to_aos()))
    else:
        st = core.FusedState(emu, c.np, path)
        st.load(c)
        for _ in range(nsteps):
            st.evolve(1)
            st.store(c)
            snaps.append(snap(c.to_aos()))
        st.info()
    return np.array(snaps)


@pytest.mark.parametrize("path", ["integrator", "fp64", "fp32"])
def test_kat_fast_flavor_gpu(core, path):
    s = run_kat(core, "inputs_fast_flavor", 100, path, ("time", "f01_Re", "f01_Im", "f01_Rebar", "f01_Imbar"))
    errs = kat.fast_flavor_errors(*s.T)
    assert max(errs) < 2e-2, errs


@pytest.mark.parametrize("path", ["integrator", "fp64", "fp32"])
def test_kat_fast_flavor_nonzerok_gpu(core, path):
    s = run_kat(core, "inputs_fast_flavor_nonzerok", 160, path,
                ("time", "f01_Re", "f01_Im", "f01_Rebar", "f01_Imbar"), reduce="max")
    errs = kat.fast_flavor_k_errors(*s.T)
    assert max(errs) < 1e-2, errs


@pytest.mark.parametrize("path", ["integrator", "fp64", "fp32"])
def test_kat_msw_gpu(core, path):
    s = run_kat(core, "inputs_msw_test", 50, path, ("time", "f00_Re", "f11_Re", "f00_Rebar", "f11_Rebar"))
    e = kat.msw_errors(*s.T)
    assert max(e.values()) < 7e-2, e
