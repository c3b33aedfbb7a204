"""emu_deposit_to_mesh_static (the deposit that allocates nothing) against
emu_deposit_to_mesh on the same inputs, through the C ABI: two calls in a row
on one GPU, the second with fewer particles on a smaller grid of another
context (the static scratch is shared by every call on the GPU, and keeps the
first call's contents), each within 1e-12 of the allocating entry point's fab
(the FP64 atomics' order differs between any two runs); and a call beyond the
scratch's capacity is refused before anything runs."""
import ctypes

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


def deposit_both(core, oracle, ncell, nphi, nkeep, seed):
    from emu_amd.capi import lib
    d = cases.fast_flavor_dict(nf=3, ncell=ncell, nphi=nphi, st4_amplitude=1e-3, seed=seed)
    deck = cases.oracle_deck(oracle, d)
    emu = core.Emu(cases.product_params(d))
    parts = oracle.init_particles(deck)
    rng = np.random.default_rng(seed)
    parts = parts[rng.permutation(len(parts))[:nkeep]]
    c = core.FlavoredNeutrinoContainer.from_aos(emu, parts)
    want = core.MultiFab(emu)
    core.deposit_to_mesh(c, want)
    got = core.MultiFab(emu)
    got.data.fill_(7.0)  # every deposited component must be overwritten
    got.data[0:3] = want.data[0:3]
    pv, gv = c.view(), got.view()
    assert lib().emu_deposit_to_mesh_static(emu.ctx, ctypes.byref(pv), ctypes.byref(gv)) == 0
    emu.sync()
    a = got.data.cpu().numpy().reshape(emu.ncomp, -1).T
    b = want.data.cpu().numpy().reshape(emu.ncomp, -1).T
    return cases.max_rel_err(a, b), len(parts)


def test_static_deposit_matches_allocating_deposit(core, oracle):
    e1, n1 = deposit_both(core, oracle, (6, 5, 4), 16, 9000, 3)
    e2, n2 = deposit_both(core, oracle, (3, 4, 2), 8, 500, 4)
    assert n2 < n1
    assert e1 < 1e-12 and e2 < 1e-12, (e1, e2)


def test_static_deposit_refuses_beyond_capacity(core):
    """A view claiming more particles than the scratch holds is refused on the
    host, before any kernel reads it."""
    from emu_amd.capi import lib
    from emu_amd import configs
    emu = core.Emu(configs.scaled(configs.cpu16(), (2, 2, 2)))
    c = core.FlavoredNeutrinoContainer(emu, 4)
    pv = c.view()
    pv.np = (1 << 22) + 1
    state = core.MultiFab(emu)
    gv = state.view()
    code = lib().emu_deposit_to_mesh_static(emu.ctx, ctypes.byref(pv), ctypes.byref(gv))
    assert code == 12
    assert b"at most 4194304 particles" in lib().emu_last_error(emu.ctx)
