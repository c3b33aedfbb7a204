"""The GPU parity tests compare several fused FP64 steps with the oracle at
1e-12 of each component's range.  That bar only means something where the
step itself does not amplify rounding: these CPU tests perturb the oracle's
initial flavour state by one ulp and measure how far the result moves after
the same steps, for the configurations of those GPU tests.  The adaptive
step at a speedup of 1e9 is far past the conservative flavour bound by
design; its measured amplification is what its own, wider bar is built on."""
import pytest

import cases

FIXED = 0.3 * 0.5 / 2.99792458e10  # the GPU tests' cell-crossing dt (0.5 cm cells)


def amplification(oracle, nf, d, nsteps, fixed_dt, parts_fn=None):
    """max over components of |x' - x| / range(x) after nsteps, where x' starts
    from the flavour state (every f and fbar component) scaled by 1 + 2^-52."""
    deck = cases.oracle_deck(oracle, d)
    prm = deck.params()
    base = oracle.init_particles(deck) if parts_fn is None else parts_fn(deck)

    def run(parts):
        g = oracle.Grid(nf, prm)
        vv = oracle.vvac_max(nf, prm, parts)
        oracle.deposit_to_mesh(nf, prm, parts, g)
        err, dt = oracle.compute_dt(nf, prm, g, vv)
        assert err == 0
        for _ in range(nsteps):
            err, dt, _ = oracle.rk4_step(nf, prm, parts, g, fixed_dt or dt, vv)
            assert err == 0
        return parts

    a = run(base.copy())
    p = base.copy()
    fcols = [i for i, n in enumerate(oracle.attribute_names(nf)) if n.startswith("f")]
    assert len(fcols) == 2 * nf * nf
    p["rdata"][:, fcols] *= 1.0 + 2.0 ** -52
    b = run(p)
    return cases.max_rel_err(b["rdata"], a["rdata"])


@pytest.mark.parametrize("nf", [2, 3])
@pytest.mark.parametrize("ncell", [(4, 4, 4), (1, 1, 8), (3, 5, 2)])
def test_fused_parity_configs_well_conditioned(oracle, nf, ncell):
    d = cases.fast_flavor_dict(nf=nf, ncell=ncell, nphi=8)
    assert amplification(oracle, nf, d, 6, FIXED) < 1e-14


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("nf", [2, 3])
def test_lower_order_config_needs_lower_density(oracle, nf, order):
    """At the fast-flavor deck's densities the cell-crossing dt is 5x the
    flavour bound and the lower orders amplify a one-ulp change past 1e-12;
    at a tenth of them (the GPU test's setting) they do not."""
    kw = dict(nf=nf, ncell=(4, 3, 5), L=(2.0, 1.5, 2.5), nphi=8, shape_order=order, st4_amplitude=1e-3)
    assert amplification(oracle, nf, cases.fast_flavor_dict(**kw), 6, FIXED) > 1e-12
    low = cases.fast_flavor_dict(**kw, st4_ndens=4.89e31, st4_ndensbar=4.89e31)
    assert amplification(oracle, nf, low, 6, FIXED) < 1e-14


@pytest.mark.parametrize("nf", [2, 3])
def test_adaptive_speedup_amplification(oracle, nf):
    """Speedup 4: well conditioned.  Speedup 1e9: a one-ulp change grows to
    ~1e-11 in 3 steps, which the GPU test's 1e-9 state bar covers a hundred
    times over."""
    ncell = (4, 3, 5)
    kw = dict(nf=nf, ncell=ncell, L=tuple(2.0 * n for n in ncell), nphi=4, st4_amplitude=1e-3, cfl_factor=0.9)
    assert amplification(oracle, nf, cases.fast_flavor_dict(**kw, max_adaptive_speedup=4.0), 3, None) < 1e-14
    big = amplification(oracle, nf, cases.fast_flavor_dict(**kw, max_adaptive_speedup=1e9), 3, None)
    assert 1e-12 < big < 1e-10


@pytest.mark.parametrize("nf", [2, 3])
@pytest.mark.parametrize("ncell", [(4, 4, 4), (3, 5, 2), (2, 3, 7), (5, 2, 4)])
def test_fp32_shape_configs_well_conditioned(oracle, nf, ncell):
    """The boxes of the FP32-bulk domain-shape test (amplitude 1e-3)."""
    d = cases.fast_flavor_dict(nf=nf, ncell=ncell, nphi=8, st4_amplitude=1e-3)
    assert amplification(oracle, nf, d, 6, FIXED) < 1e-14


def test_ill_conditioned_box_exists(oracle):
    """5 x 2 x 3 at the same settings is not usable for a parity test."""
    d = cases.fast_flavor_dict(nf=2, ncell=(5, 2, 3), nphi=8, st4_amplitude=1e-3)
    assert amplification(oracle, 2, d, 6, FIXED) > 1e-12
