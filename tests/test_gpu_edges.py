"""Edge cases of the flavor step on the GPU, against the oracle: no particles
at all, ragged cell occupancy around the fused kernel's 64-particle chunks
(empty cells, single particles, 63/64/65 and 127/128/129 particles), and one
crowded cell that spans seven chunks; each in both slot layouts of the fused
state (cells back to back, and each cell's run on a 128-byte boundary), and
the two layouts against each other."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def core():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from emu_amd import core as c
    return c


def make(core, oracle, d):
    deck = cases.oracle_deck(oracle, d)
    return deck, deck.params(), core.Emu(cases.product_params(d))


def placed(oracle, deck, ncell, L, occupancy, seed):
    """Valid particles of the deck's initial conditions (normalised flavour
    states, their lengths, weights and directions), moved into the cells of
    `occupancy` (linear cell index -> count) at uniform positions inside each."""
    src = oracle.init_particles(deck)
    n = sum(occupancy.values())
    assert n <= len(src)
    parts = src[:n].copy()
    rng = np.random.default_rng(seed)
    dx = np.array(L) / np.array(ncell)
    i = 0
    for cell, cnt in sorted(occupancy.items()):
        cx, cy, cz = cell % ncell[0], (cell // ncell[0]) % ncell[1], cell // (ncell[0] * ncell[1])
        for _ in range(cnt):
            pos = (np.array([cx, cy, cz]) + rng.uniform(0.05, 0.95, 3)) * dx
            parts["pos"][i] = pos
            parts["rdata"][i, 1:4] = pos
            i += 1
    parts["id"] = np.arange(1, n + 1)
    return parts


def run_fused_against_oracle(core, oracle, nf, prm, emu, parts, prec, nsteps, fixed, slots):
    c = core.FlavoredNeutrinoContainer.from_aos(emu, parts)
    st = core.FusedState(emu, c.np, prec, slots)
    st.load(c)
    st.evolve(nsteps, fixed_dt=fixed)
    st.store(c)
    info = st.info()
    st.close()
    g = oracle.Grid(nf, prm)
    vv = oracle.vvac_max(nf, prm, parts)
    oracle.deposit_to_mesh(nf, prm, parts, g)
    for _ in range(nsteps):
        err, _, bad = oracle.rk4_step(nf, prm, parts, g, fixed, vv)
        assert err == 0
    got = cases.by_id(c.to_aos())
    want = cases.by_id(parts)
    np.testing.assert_array_equal(got["id"], want["id"])
    if prec == "fp64":
        np.testing.assert_array_equal(got["pos"], want["pos"])
        assert cases.max_rel_err(got["rdata"], want["rdata"]) < 1e-12
    else:
        # FP32 bulk keeps the momenta in FP32: the FP64 position update moves
        # along a direction rounded to ~2^-23, so the position error is bounded
        # by that fraction of the distance travelled (absolute, not relative:
        # coordinates near 0 have no relative bound)
        travel = nsteps * fixed * oracle.clight
        assert np.abs(got["pos"] - want["pos"]).max() <= 2.0 ** -22 * travel
        assert cases.max_rel_err(got["rdata"], want["rdata"]) < 2e-5
    assert info["steps"] == nsteps


@pytest.mark.parametrize("slots", ["packed", "aligned"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("nf", [2, 3])
def test_ragged_occupancy(core, oracle, nf, prec, slots):
    """Cells holding 0, 1, 2, 63, 64, 65, 127, 128, 129, 200, 17, 5 or 40
    particles, 4 steps with a dt that moves particles across cells."""
    ncell, L = (4, 4, 4), (2.0, 2.0, 2.0)
    d = cases.fast_flavor_dict(nf=nf, ncell=ncell, L=L, nphi=16, st4_amplitude=1e-3)
    deck, prm, emu = make(core, oracle, d)
    counts = [0, 1, 2, 63, 64, 65, 127, 128, 129, 200, 0, 0, 17, 5, 0, 40]
    occ = {cell: counts[cell % len(counts)] for cell in range(64)}
    parts = placed(oracle, deck, ncell, L, occ, seed=5)
    run_fused_against_oracle(core, oracle, nf, prm, emu, parts, prec, 4, 0.3 * 0.5 / oracle.clight, slots)


@pytest.mark.parametrize("slots", ["packed", "aligned"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_crowded_cell(core, oracle, prec, slots):
    """400 particles in one cell (six full chunks and a partial one), its
    neighbours empty, and a few particles elsewhere."""
    nf, ncell, L = 3, (3, 3, 3), (1.5, 1.5, 1.5)
    d = cases.fast_flavor_dict(nf=nf, ncell=ncell, L=L, nphi=16, st4_amplitude=1e-3)
    deck, prm, emu = make(core, oracle, d)
    parts = placed(oracle, deck, ncell, L, {13: 400, 0: 3, 26: 1}, seed=9)
    run_fused_against_oracle(core, oracle, nf, prm, emu, parts, prec, 3, 0.3 * 0.5 / oracle.clight, slots)


@pytest.mark.parametrize("slots", ["packed", "aligned"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_no_particles(core, oracle, prec, slots):
    """An empty container: the deposit leaves only the background fields, the
    flavour bound of compute_dt has nothing to bound (the step is the
    translation bound, as in the reference's compute_dt), and the fused state
    steps through time with nothing to move."""
    nf, ncell, L = 2, (3, 3, 3), (1.5, 1.5, 1.5)
    d = cases.fast_flavor_dict(nf=nf, ncell=ncell, L=L, nphi=4)
    deck, prm, emu = make(core, oracle, d)
    empty = oracle.init_particles(deck)[:0]
    c = core.FlavoredNeutrinoContainer.from_aos(emu, empty)
    assert c.np == 0
    state = core.MultiFab(emu)
    core.deposit_to_mesh(c, state)
    state.FillBoundary()
    g = oracle.Grid(nf, prm)
    oracle.deposit_to_mesh(nf, prm, empty, g)
    oracle.fill_boundary(nf, prm, g)
    np.testing.assert_array_equal(state.data.cpu().numpy().reshape(-1), g.fab().reshape(-1))
    err, dt_or = oracle.compute_dt(nf, prm, g, 0.0)
    assert err == 0 and dt_or == pytest.approx(0.4 * 0.5 / oracle.clight, rel=1e-14)
    assert core.compute_dt(state, 0.0) == pytest.approx(dt_or, rel=1e-14)
    c.Renormalize()
    st = core.FusedState(emu, 0, prec, slots)
    st.load(c)
    st.evolve(2)
    info = st.info()
    st.store(c)
    st.close()
    assert info["steps"] == 2
    assert info["time"] == pytest.approx(2 * dt_or, rel=1e-14)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("nf", [2, 3])
def test_slot_layouts_agree(core, oracle, nf, prec):
    """The aligned layout only moves where each cell's run of particles starts;
    the order within a cell is that of the packed layout.  What still differs
    between any two runs is the order in which the deposit's FP64 atomics
    from different cells land, at the grid's last bit.  So: 6 steps with a
    dt that moves particles across cells give the same ids and positions bit
    for bit and flavour states within that rounding; then 2 steps with
    compute_dt's dt (the auxiliary deposit) give the same dt to rounding.  A
    6 x 5 x 4 box of 92 directions: runs of 92, and their neighbours after
    moves."""
    d = cases.fast_flavor_dict(nf=nf, ncell=(6, 5, 4), nphi=16, st4_amplitude=1e-3)
    deck, prm, emu = make(core, oracle, d)
    parts = oracle.init_particles(deck)
    out = {}
    for slots in ("packed", "aligned"):
        c = core.FlavoredNeutrinoContainer.from_aos(emu, parts)
        st = core.FusedState(emu, c.np, prec, slots)
        st.load(c)
        st.evolve(6, fixed_dt=0.3 * 0.5 / oracle.clight)
        st.store(c)
        fixed = c.to_aos()
        st.evolve(2)
        out[slots] = (fixed, st.info())
        st.close()
    (a, ia), (b, ib) = out["packed"], out["aligned"]
    np.testing.assert_array_equal(a["id"], b["id"])
    np.testing.assert_array_equal(a["pos"], b["pos"])
    assert cases.max_rel_err(b["rdata"], a["rdata"]) < (1e-13 if prec == "fp64" else 2e-5)
    assert ib["steps"] == ia["steps"] == 8
    assert ib["dt"] == pytest.approx(ia["dt"], rel=1e-12)
    assert ib["time"] == pytest.approx(ia["time"], rel=1e-12)


@pytest.mark.parametrize("nf", [2, 3])
@pytest.mark.parametrize("ncell", [(4, 4, 4), (3, 5, 2), (2, 3, 7), (5, 2, 4)])
def test_fp32_bulk_domain_shapes(core, oracle, nf, ncell):
    """FP32-bulk takes two x-adjacent cells per wave: an odd x count leaves the
    last column's wave one cell, and a 2-cell x dimension makes the pair's
    4-wide box hold each periodic image twice.  6 steps with a dt that moves
    particles across cells, against the oracle.  (The boxes are ones where the
    oracle's own steps keep a one-ulp change at 1e-15: 5 x 2 x 3, for one,
    grows it to 7e-11, which FP32's rounding would turn into 1e-2.)"""
    d = cases.fast_flavor_dict(nf=nf, ncell=ncell, nphi=8, st4_amplitude=1e-3)
    deck, prm, emu = make(core, oracle, d)
    parts = oracle.init_particles(deck)
    run_fused_against_oracle(core, oracle, nf, prm, emu, parts, "fp32", 6, 0.3 * 0.5 / oracle.clight, "aligned")


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("nf", [2, 3])
def test_fp32_bulk_lower_shape_orders(core, oracle, nf, order):
    """FP32-bulk at SHAPE_FACTOR_ORDER 0 (1-point stencils: one cell per wave)
    and 1 (3-point stencils with a zero weight: two cells per wave), 6 steps
    with a dt that moves particles across cells, against the oracle at the same
    order.  A tenth of the fast-flavor densities, where this dt is within the
    flavour bound and the step is well conditioned at these orders."""
    d = cases.fast_flavor_dict(nf=nf, ncell=(4, 3, 5), L=(2.0, 1.5, 2.5), nphi=8, shape_order=order,
                               st4_amplitude=1e-3, st4_ndens=4.89e31, st4_ndensbar=4.89e31)
    deck, prm, emu = make(core, oracle, d)
    parts = oracle.init_particles(deck)
    run_fused_against_oracle(core, oracle, nf, prm, emu, parts, "fp32", 6, 0.3 * 0.5 / oracle.clight, "aligned")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("kind", ["trace", "negdiag", "length", "move"])
def test_fused_deferred_errors(core, oracle, kind, prec):
    """The fused step defers its errors to emu_state_info, which names the
    lowest offending particle id: Renormalize's three aborts
    (generate_code.py:404-409, 418-423, 433-438 inside
    FlavoredNeutrinoContainer.cpp:89-113) at the end of the step, and a
    particle that outruns the one-cell move the cell sort allows (only
    |p| > p_t can: here p_x = 3 p_t with c dt = 0.9 dx, a move of 2 or 3
    cells along x in a 6-cell dimension).  Two particles are made bad, the
    higher id in the lower slot, so the lowest id has to win."""
    nf, ncell, L = 2, (6, 4, 4), (3.0, 2.0, 2.0)
    d = cases.fast_flavor_dict(nf=nf, ncell=ncell, L=L, nphi=4)
    deck, prm, emu = make(core, oracle, d)
    parts = placed(oracle, deck, ncell, L, {c: 6 for c in range(96)}, seed=3)
    K = {nm: i for i, nm in enumerate(oracle.attribute_names(nf))}
    bad_ids = (41, 17)  # slots 40 and 16
    want_code = {"trace": 1, "negdiag": 2, "length": 3, "move": 14}[kind]
    for bid in bad_ids:
        r = parts["rdata"][bid - 1]
        if kind == "trace":
            r[K["f00_Re"]] += 1e-3
        elif kind == "negdiag":
            r[K["f00_Re"]], r[K["f11_Re"]] = -0.01, 1.01
        elif kind == "length":
            r[K["L"]] *= 2.0
        else:
            r[K["pupx"]], r[K["pupy"]], r[K["pupz"]] = 3.0 * r[K["pupt"]], 0.0, 0.0
    c = core.FlavoredNeutrinoContainer.from_aos(emu, parts)
    st = core.FusedState(emu, c.np, prec, "aligned")
    st.load(c)
    assert st.info()["code"] == 0
    st.evolve(1, fixed_dt=0.9 * 0.5 / oracle.clight)
    info = st.info(raise_on_error=False)
    with pytest.raises(Exception) as ei:
        st.info()
    assert ei.value.code == want_code and ei.value.bad_index == min(bad_ids)
    st.store(c)
    st.close()
    assert info["code"] == want_code, info
    assert info["bad_index"] == min(bad_ids), info
    assert info["steps"] == 1
