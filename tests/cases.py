"""Shared helpers: one parameter dict drives both the oracle (pyoracle.Deck)
and the product (emu_amd TestParams), so both see the same inputs."""
import math

import numpy as np

import kat


def fast_flavor_dict(nf=3, ncell=(4, 4, 4), L=None, directions="sphere", nphi=4, **kw):
    L = L if L is not None else tuple(0.5 * n for n in ncell)
    d = dict(nf=nf, simulation_type=4, ncell=tuple(ncell), Lx=L[0], Ly=L[1], Lz=L[2], nphi_equator=nphi,
             directions=directions, st4_ndens=4.89e32, st4_ndensbar=4.89e32, st4_fluxfac=0.0,
             st4_fluxfacbar=0.33, st4_theta=0.0, st4_thetabar=0.0, st4_amplitude=1e-6,
             cfl_factor=0.4, flavor_cfl_factor=0.4, maxError=1e-6, rho_g_ccm=0.0, Ye=1.0,
             mass1_eV=0.0, mass2_eV=math.sqrt(7.4e-5), mass3_eV=math.sqrt(2.5e-3),
             theta12_degrees=33.82, theta13_degrees=48.3, theta23_degrees=8.61, deltaCP_degrees=222.0, seed=99)
    if nf == 2:
        d.update(mass1_eV=-math.sqrt(7.4e-5), mass2_eV=0.0, mass3_eV=0.0)
    d.update(kw)
    return d


def deck_dict(name, nf=2):
    d = kat.reference_deck(name)
    d = {k: (tuple(v) if isinstance(v, list) else v) for k, v in d.items()}
    d["nf"] = nf
    return d


def oracle_deck(oracle, d):
    fields = oracle.Deck.__dataclass_fields__
    kw = {k: v for k, v in d.items() if k in fields}
    st5 = {k[4:]: v for k, v in d.items() if k.startswith("st5_")}
    if st5:
        kw["st5"] = st5
    return oracle.Deck(**kw)


def minerbo_dict(nf=3, ncell=(2, 2, 2), nphi=32, **kw):
    """Simulation type 5 (the Minerbo closure, FlavoredNeutrinoContainerInit.cpp
    :574-654): nu_e isotropic, nubar_e with flux factor 0.33 along +z, heavy
    leptons with 0.1 along a tilted axis."""
    d = fast_flavor_dict(nf=nf, ncell=ncell, nphi=nphi, simulation_type=5)
    d.update(st5_nnue=4.89e32, st5_nnua=4.0e32, st5_nnux=2.0e32, st5_fxnue=0.0, st5_fynue=0.0, st5_fznue=0.0,
             st5_fxnua=0.0, st5_fynua=0.0, st5_fznua=0.33, st5_fxnux=0.06, st5_fynux=-0.048, st5_fznux=0.064,
             st5_amplitude=1e-6)
    d.update(kw)
    return d


def product_params(d):
    from emu_amd.params import TestParams
    return TestParams.from_dict(d)


def by_id(arr):
    return arr[np.argsort(arr["id"], kind="stable")]


def max_rel_err(a, b):
    """max over components of |a - b| / max|b| (per-component normalisation:
    per-element relative error is ill-defined for cancelled fluxes)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.abs(b).max(axis=0)
    scale = np.where(scale > 0, scale, 1.0)
    return float((np.abs(a - b) / scale).max())


# keys Parameters.H:53-130 requires, with the values a deck dict may leave out
_DECK_DEFAULTS = dict(nppc=(1, 1, 1), max_grid_size=64, nsteps=1, end_time=1.0, T_MeV=10.0, max_adaptive_speedup=0.0,
                      write_plot_every=0, write_plot_particles_every=0, do_restart=0, restart_dir="",
                      alpha1_degrees=0.0, alpha2_degrees=0.0, st4_phi=0.0, st4_phibar=0.0)


def deck_text(d: dict) -> str:
    """A ParmParse inputs file (key = value) for a deck dict."""
    full = {**_DECK_DEFAULTS, **d}
    lines = []
    for k, v in full.items():
        if isinstance(v, (tuple, list)):
            v = "(" + ",".join(str(int(x)) for x in v) + ")"
        elif isinstance(v, str):
            v = f'"{v}"'
        elif isinstance(v, float):
            v = repr(v)
        lines.append(f"{k} = {v}")
    return "\n".join(lines) + "\n"
