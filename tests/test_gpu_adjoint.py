"""The hand-written adjoint: forward RHS against the oracle, the VJP by the
dot-product test, and whole-run gradients against central finite differences
(1e-5 relative) through torch.autograd."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    import torch
    assert torch.cuda.is_available()
    from emu_amd import adjoint, core
    return torch, core, adjoint


def setup(mods, oracle, nf, ncell=(4, 4, 4), nphi=4, amp=1e-4, **kw):
    torch, core, adjoint = mods
    d = cases.fast_flavor_dict(nf=nf, ncell=ncell, nphi=nphi, st4_amplitude=amp, maxError=1e-2, **kw)
    deck = cases.oracle_deck(oracle, d)
    parts = oracle.init_particles(deck)
    emu = core.Emu(cases.product_params(d))
    c = core.FlavoredNeutrinoContainer.from_aos(emu, parts)
    return emu, c, deck, parts


@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("nf", [2, 3])
def test_adj_rhs_matches_oracle(mods, oracle, nf, order):
    """The adjoint's forward RHS (cell-listed MFMA deposit, unrolled gather) at
    each shape order against the oracle's deposit + interpolation."""
    torch, core, adjoint = mods
    emu, c, deck, parts = setup(mods, oracle, nf, rho_g_ccm=1e9, Ye=0.5, shape_order=order)
    prob = adjoint.AdjointProblem(emu, c, dt=1e-13, nsteps=1)
    k = prob.rhs(prob.pos0, prob.Nn_base, prob.f0).cpu().numpy()
    prm = deck.params()
    g = oracle.Grid(nf, prm)
    oracle.deposit_to_mesh(nf, prm, parts, g)
    oracle.fill_boundary(nf, prm, g)
    rhs = parts.copy()
    oracle.interpolate_rhs_from_mesh(nf, prm, rhs, g)
    K = emu.K
    nh = nf * nf
    want = np.concatenate([rhs["rdata"][:, K["f00_Re"]:K["f00_Re"] + nh],
                           rhs["rdata"][:, K["f00_Rebar"]:K["f00_Rebar"] + nh]], axis=1)
    assert cases.max_rel_err(k.T, want) < 1e-12


@pytest.mark.parametrize("nf", [2, 3])
def test_vjp_dot_product(mods, oracle, nf):
    """F is quadratic in Y (H is affine in Y, the commutator bilinear), so the
    central difference of lam . F is exact up to roundoff."""
    torch, core, adjoint = mods
    emu, c, deck, parts = setup(mods, oracle, nf, amp=1e-2)
    prob = adjoint.AdjointProblem(emu, c, dt=1e-13, nsteps=1)
    gen = torch.Generator(device="cpu").manual_seed(3)
    Y = prob.f0
    lam = torch.randn(Y.shape, generator=gen, dtype=torch.float64).to(Y.device)
    dY = torch.randn(Y.shape, generator=gen, dtype=torch.float64).to(Y.device) * 1e-3
    Nn = prob.Nn_base
    prob.lam_nbar.zero_()
    lamY = prob.vjp(prob.pos0, Nn, Y, lam)
    h = 1.0
    fp = prob.rhs(prob.pos0, Nn, Y + h * dY)
    fm = prob.rhs(prob.pos0, Nn, Y - h * dY)
    lhs = float((lam * (fp - fm)).sum() / (2 * h))
    rhs_ = float((lamY * dY).sum())
    assert rhs_ == pytest.approx(lhs, rel=1e-9)
    # Nbar direction
    dN = torch.zeros_like(Nn)
    dN[1] = torch.rand(Nn.shape[1], generator=gen, dtype=torch.float64).to(Y.device) * Nn[1] * 1e-3
    fp = prob.rhs(prob.pos0, Nn + dN, Y)
    fm = prob.rhs(prob.pos0, Nn - dN, Y)
    lhs = float((lam * (fp - fm)).sum() / 2)
    rhs_ = float((prob.lam_nbar * dN[1]).sum())
    assert rhs_ == pytest.approx(lhs, rel=1e-6)


@pytest.mark.parametrize("nf", [2, 3])
def test_adjoint_gradient_matches_finite_differences(mods, oracle, nf):
    torch, core, adjoint = mods
    emu, c, deck, parts = setup(mods, oracle, nf, amp=1e-6)
    dt = 0.4 * 0.5 / 2.99792458e10 * 0.25
    prob = adjoint.AdjointProblem(emu, c, dt=dt, nsteps=8, checkpoint_every=3)
    f0 = prob.f0.clone().requires_grad_(True)
    ratio = torch.tensor(1.0, dtype=torch.float64, device=f0.device, requires_grad=True)
    J = adjoint.FlavorObjective.apply(f0, ratio, prob)
    J.backward()
    g = f0.grad
    nh = nf * nf
    K = emu.K
    i01 = K["f01_Re"] - K["f00_Re"]
    rng = np.random.default_rng(0)
    picks = [(i01, p) for p in rng.choice(c.np, 4, replace=False)] + \
            [(i01 + 1, p) for p in rng.choice(c.np, 2, replace=False)] + \
            [(nh + i01, p) for p in rng.choice(c.np, 2, replace=False)]
    for comp, p in picks:
        base = float(prob.f0[comp, p])
        h = 1e-4 * abs(base) if base != 0 else 1e-10
        fp = prob.f0.clone()
        fp[comp, p] += h
        fm = prob.f0.clone()
        fm[comp, p] -= h
        Jp, _, _, _, _ = prob.run(fp, 1.0, keep_checkpoints=False)
        Jm, _, _, _, _ = prob.run(fm, 1.0, keep_checkpoints=False)
        fd = (Jp - Jm) / (2 * h)
        assert float(g[comp, p]) == pytest.approx(fd, rel=1e-5, abs=1e-5 * abs(fd) + 1e-300), (comp, p)
    h = 1e-4
    Jp, _, _, _, _ = prob.run(prob.f0, 1.0 + h, keep_checkpoints=False)
    Jm, _, _, _, _ = prob.run(prob.f0, 1.0 - h, keep_checkpoints=False)
    fd = (Jp - Jm) / (2 * h)
    assert float(ratio.grad) == pytest.approx(fd, rel=1e-5)


@pytest.mark.parametrize("order", [0, 1, 2])
@pytest.mark.parametrize("nf", [2, 3])
def test_fused_step_matches_per_rhs_step(mods, oracle, nf, order):
    """emu_adj_step (one call: the stages with their deposits, the
    accumulator, Renormalize, the positions) against the same step composed
    from the per-RHS calls and axpys, over 3 steps; and its VJP against the
    central difference of <lam, step(y)> along a random direction (the step is
    a polynomial in y, so the O(h^2) error is far below the bar).  At each
    shape order: the step kernels deposit onto the full 3 x 3 x 3
    neighbourhood with the order's weights."""
    torch, core, adjoint = mods
    emu, c, deck, parts = setup(mods, oracle, nf, amp=1e-3, shape_order=order)
    dt = 0.4 * 0.5 / 2.99792458e10 * 0.5
    prob = adjoint.AdjointProblem(emu, c, dt=dt, nsteps=3)
    pa, pb = prob.pos0.clone(), prob.pos0.clone()
    ya, yb = prob.f0.clone(), prob.f0.clone()
    for _ in range(3):
        ya = prob.step(pa, prob.Nn_base, ya)
        yb = prob.step_unfused(pb, prob.Nn_base, yb)
    prob.check_status()
    assert torch.equal(pa, pb)
    scale = yb.abs().amax(dim=1, keepdim=True).clamp_min(1e-300)
    assert float(((ya - yb).abs() / scale).max()) < 1e-12
    gen = torch.Generator(device="cpu").manual_seed(5)
    y = prob.f0
    lam = torch.randn(y.shape, generator=gen, dtype=torch.float64).to(y.device)
    dy = torch.randn(y.shape, generator=gen, dtype=torch.float64).to(y.device) * 1e-4
    prob.lam_nbar.zero_()
    g = prob.step_adjoint(prob.pos0, prob.Nn_base, y, lam)
    h = 1e-2

    def J(yy):
        p = prob.pos0.clone()
        return prob.step(p, prob.Nn_base, yy)
    fd = float((lam * (J(y + h * dy) - J(y - h * dy))).sum() / (2 * h))
    assert float((g * dy).sum()) == pytest.approx(fd, rel=1e-8)


@pytest.mark.parametrize("nf", [2, 3])
def test_kept_stages_match_recompute(mods, oracle, nf):
    """emu_adj_step_keep gives emu_adj_step's y_{n+1} and positions, and
    emu_adj_step_vjp_kept (the stage states and deposits taken from the kept
    buffer) the recomputing VJP's lam_n and lam_Nbar; the four segment modes
    give one gradient.  The deposits' atomics add in no set order, so the bar
    is 1e-12 of each array's largest entry, not equality."""
    torch, core, adjoint = mods
    emu, c, deck, parts = setup(mods, oracle, nf, amp=1e-3)
    dt = 0.4 * 0.5 / 2.99792458e10 * 0.5
    prob = adjoint.AdjointProblem(emu, c, dt=dt, nsteps=3)
    keep = torch.empty(prob.keep_doubles, dtype=torch.float64, device=prob.f0.device)

    def close(a, b):
        return float((a - b).abs().max() / b.abs().max().clamp_min(1e-300)) < 1e-12

    pa, pb = prob.pos0.clone(), prob.pos0.clone()
    ya = prob.step(pa, prob.Nn_base, prob.f0)
    yb = prob.step(pb, prob.Nn_base, prob.f0, keep=keep)
    prob.check_status()
    assert torch.equal(pa, pb)
    assert close(yb, ya)
    gen = torch.Generator(device="cpu").manual_seed(7)
    lam = torch.randn(prob.f0.shape, generator=gen, dtype=torch.float64).to(prob.f0.device)
    prob.lam_nbar.zero_()
    la = prob.step_adjoint(prob.pos0, prob.Nn_base, prob.f0, lam)
    na = prob.lam_nbar.clone()
    prob.lam_nbar.zero_()
    lb = prob.step_adjoint(prob.pos0, prob.Nn_base, prob.f0, lam, keep=keep)
    assert close(lb, la)
    assert close(prob.lam_nbar, na)
    grads = {}
    for seg in ("keep", "store", "recompute", "binomial"):
        p = adjoint.AdjointProblem(emu, c, dt=dt, nsteps=8, checkpoint_every=3, segment=seg)
        grads[seg] = p.gradients(p.f0.clone(), 1.0)
    # the binomial schedule over one 8-step segment with one snapshot: its
    # snapshots fall between the re-sorts (every 4 steps), and it runs the
    # fewest forward steps revolve allows (8 for the run, 14 replayed)
    p = adjoint.AdjointProblem(emu, c, dt=dt, nsteps=8, checkpoint_every=8, segment="binomial", snapshots=1)
    grads["binomial1"] = p.gradients(p.f0.clone(), 1.0)
    assert p.steps_run == 8 + 14
    for seg in ("store", "recompute", "binomial", "binomial1"):
        assert grads[seg][0] == pytest.approx(grads["keep"][0], rel=1e-12)
        assert close(grads[seg][1], grads["keep"][1])
        assert grads[seg][2] == pytest.approx(grads["keep"][2], rel=1e-10)


@pytest.mark.parametrize("nf", [2, 3])
def test_fp32_segment_modes_agree(mods, oracle, nf):
    """The FP32-bulk gradient through the four segment modes (kept stage
    data, kept step states, recompute from the checkpoint, the binomial
    schedule), 8 steps with checkpoints every 3, and the binomial schedule
    over one 8-step segment: the same J and gradients to FP32 rounding (the FP64
    deposit sums add in no set order, which can move an FP32 state by an
    ulp)."""
    torch, core, adjoint = mods
    emu, c, deck, parts = setup(mods, oracle, nf, amp=1e-3)
    dt = 0.4 * 0.5 / 2.99792458e10 * 0.5
    grads = {}
    for seg in ("keep", "store", "recompute", "binomial"):
        p = adjoint.AdjointProblem(emu, c, dt=dt, nsteps=8, checkpoint_every=3, segment=seg, precision="fp32")
        grads[seg] = p.gradients(p.f0.clone(), 1.0)
    p = adjoint.AdjointProblem(emu, c, dt=dt, nsteps=8, checkpoint_every=8, segment="binomial", snapshots=1,
                               precision="fp32")
    grads["binomial1"] = p.gradients(p.f0.clone(), 1.0)
    J, g, r = grads["keep"]
    assert g.dtype == torch.float64
    for seg in ("store", "recompute", "binomial", "binomial1"):
        J2, g2, r2 = grads[seg]
        assert J2 == pytest.approx(J, rel=1e-6)
        assert float((g2 - g).abs().max() / g.abs().max()) < 1e-5
        assert r2 == pytest.approx(r, rel=1e-5)


def test_adjoint32_gradient_matches_finite_differences(mods):
    """The two-flavour 32^3 x 92 workload (3,014,656 particles), 200 RK4
    steps with checkpoints every 20: dJ/dratio and directional derivatives of
    J along two random directions of the initial flavour state, each against
    a central difference with h = 1e-4 (relative to the ratio, and to the
    direction's own scale of the 1e-6 perturbations), to 1e-5."""
    torch, core, adjoint = mods
    from emu_amd import configs
    parms = configs.adjoint32()
    emu = core.Emu(parms)
    c = core.FlavoredNeutrinoContainer.init_particles(emu)
    assert c.np == 3014656
    st = core.FusedState(emu, c.np, "fp64")
    st.load(c)
    dt = st.info()["dt"]
    st.close()
    prob = adjoint.AdjointProblem(emu, c, dt=dt, nsteps=200, checkpoint_every=20)
    J0, gf, gr = prob.gradients(prob.f0.clone(), 1.0)
    h = 1e-4
    Jp, _, _, _, _ = prob.run(prob.f0, 1.0 + h, keep_checkpoints=False)
    Jm, _, _, _, _ = prob.run(prob.f0, 1.0 - h, keep_checkpoints=False)
    fd = (Jp - Jm) / (2 * h)
    print(f"adjoint32: J {J0:.6e}, dJ/dratio adjoint {gr:.10e} fd {fd:.10e}")
    assert gr == pytest.approx(fd, rel=1e-5)
    gen = torch.Generator(device="cpu").manual_seed(11)
    K = emu.K
    nh = emu.nh
    i01 = K["f01_Re"] - K["f00_Re"]
    for trial in range(2):
        v = torch.zeros_like(prob.f0)
        # the off-diagonal components the perturbation lives in, nu and nubar
        for comp in (i01, i01 + 1, nh + i01, nh + i01 + 1):
            v[comp] = torch.randn(c.np, generator=gen, dtype=torch.float64).to(v.device) * 1e-6
        Jp, _, _, _, _ = prob.run(prob.f0 + h * v, 1.0, keep_checkpoints=False)
        Jm, _, _, _, _ = prob.run(prob.f0 - h * v, 1.0, keep_checkpoints=False)
        fd = (Jp - Jm) / (2 * h)
        ad = float((gf * v).sum())
        print(f"adjoint32: direction {trial}: adjoint {ad:.10e} fd {fd:.10e}")
        assert ad == pytest.approx(fd, rel=1e-5)


@pytest.mark.parametrize("nf", [2, 3])
def test_fp32_step_and_vjp_track_fp64(mods, oracle, nf):
    """The FP32-bulk step calls against the FP64 ones from the same state: one
    step's y_{n+1} and the step VJP's lam_n and lam_Nbar agree to FP32
    rounding (1e-5 of each array's largest entry), and the FP32 positions are
    the FP64 ones exactly (positions stay FP64)."""
    torch, core, adjoint = mods
    emu, c, deck, parts = setup(mods, oracle, nf, amp=1e-3)
    dt = 0.4 * 0.5 / 2.99792458e10 * 0.5
    p64 = adjoint.AdjointProblem(emu, c, dt=dt, nsteps=3)
    p32 = adjoint.AdjointProblem(emu, c, dt=dt, nsteps=3, precision="fp32")

    def close(a, b, tol):
        return float((a.double() - b.double()).abs().max() / b.double().abs().max().clamp_min(1e-300)) < tol

    pa, pb = p64.pos0.clone(), p32.pos0.clone()
    ya = p64.step(pa, p64.Nn_base, p64.f0)
    yb = p32.step(pb, p32.Nn_base, p32.f0)
    p64.check_status()
    p32.check_status()
    assert yb.dtype == torch.float32
    assert torch.equal(pa, pb)
    assert close(yb, ya, 1e-5)
    gen = torch.Generator(device="cpu").manual_seed(9)
    lam = torch.randn(p64.f0.shape, generator=gen, dtype=torch.float64).to(p64.f0.device)
    p64.lam_nbar.zero_()
    p32.lam_nbar.zero_()
    la = p64.step_adjoint(p64.pos0, p64.Nn_base, p64.f0, lam)
    lb = p32.step_adjoint(p32.pos0, p32.Nn_base, p32.f0, lam.float())
    assert close(lb, la, 1e-5)
    # the FP32 form's weights carry wscale, so its lam_Nbar is d/d(wscale Nbar)
    assert p32.wscale != 1.0
    assert close(p32.lam_nbar * p32.wscale, p64.lam_nbar, 1e-5)


def test_adjoint32_fp32_gradient_tracks_fp64(mods):
    """The 2-flavour 32^3 x 92 workload over 200 steps (checkpoints every 20)
    in the FP32-bulk form against the FP64 adjoint: J, dJ/dratio and the
    directional derivative of J along a random direction of the initial
    flavour state.  The bars (1e-3 relative) hold the FP32 rounding of 200
    steps of the state and its adjoint; the measured errors are printed."""
    torch, core, adjoint = mods
    from emu_amd import configs
    parms = configs.adjoint32()
    emu = core.Emu(parms)
    c = core.FlavoredNeutrinoContainer.init_particles(emu)
    st = core.FusedState(emu, c.np, "fp64")
    st.load(c)
    dt = st.info()["dt"]
    st.close()
    p64 = adjoint.AdjointProblem(emu, c, dt=dt, nsteps=200, checkpoint_every=20)
    J64, g64, r64 = p64.gradients(p64.f0.clone(), 1.0)
    del p64
    p32 = adjoint.AdjointProblem(emu, c, dt=dt, nsteps=200, checkpoint_every=20, precision="fp32")
    J32, g32, r32 = p32.gradients(p32.f0.clone(), 1.0)
    gen = torch.Generator(device="cpu").manual_seed(13)
    v = torch.randn(g64.shape, generator=gen, dtype=torch.float64).to(g64.device)
    d64 = float((g64 * v).sum())
    d32 = float((g32.double() * v).sum())
    eJ, er, ed = abs(J32 / J64 - 1), abs(r32 / r64 - 1), abs(d32 / d64 - 1)
    print(f"adjoint32 fp32 vs fp64: J {eJ:.2e}, dJ/dratio {er:.2e}, directional {ed:.2e}")
    assert eJ < 1e-3 and er < 1e-3 and ed < 1e-3



@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_footprint_bounds_the_peak(mods, prec):
    """AdjointProblem.footprint_bytes() bounds the device memory a gradient
    allocates in every segment mode, on the two-flavour 32^3 x 92 workload
    (20 steps, checkpoints every 10), and a budget below it is refused before
    anything runs."""
    torch, core, adjoint = mods
    from emu_amd import capi, configs
    emu = core.Emu(configs.adjoint32())
    c = core.FlavoredNeutrinoContainer.init_particles(emu)
    st = core.FusedState(emu, c.np, "fp64")
    st.load(c)
    dt = st.info()["dt"]
    st.close()
    del st
    for seg in ("keep", "store", "recompute", "binomial"):
        torch.cuda.synchronize()
        base = torch.cuda.memory_allocated()
        torch.cuda.reset_peak_memory_stats()
        p = adjoint.AdjointProblem(emu, c, dt=dt, nsteps=20, checkpoint_every=10, segment=seg, precision=prec,
                                   snapshots=2)
        need = p.footprint_bytes()
        p.gradients(p.f0, 1.0)
        torch.cuda.synchronize()
        peak = torch.cuda.max_memory_allocated() - base
        print(f"{prec} {seg}: peak {peak / 2**30:.3f} GiB, footprint {need / 2**30:.3f} GiB")
        assert peak <= need
        assert peak >= 0.4 * need
        del p
    with pytest.raises(capi.EmuError):
        adjoint.AdjointProblem(emu, c, dt=dt, nsteps=20, checkpoint_every=10, segment="binomial", precision=prec,
                               memory_budget=0.5 * need)
