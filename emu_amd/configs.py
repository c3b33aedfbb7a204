"""Named workloads: the five configurations the step is measured and checked on.

Each returns a TestParams (deck units).  Directions: "sphere" with
nphi_equator=16 is the reference's uniform_sphere_xyz set of 92 directions
(also used by sample_inputs/inputs_1d_fiducial:28); 32 gives 378.  The 26
directions of the small 2-flavour case are the cube's face, edge and corner
vectors.
"""
from __future__ import annotations

import math

from .params import TestParams

# fast-flavor densities of sample_inputs/inputs_1d_fiducial:2-10
NDENS = 4.89e32


def _fast_flavor(**kw) -> TestParams:
    base = dict(simulation_type=4, st4_ndens=NDENS, st4_ndensbar=NDENS, st4_theta=0.0, st4_thetabar=0.0,
                st4_phi=0.0, st4_phibar=0.0, st4_fluxfac=0.0, st4_fluxfacbar=0.33, st4_amplitude=1e-6,
                rho_g_ccm=0.0, Ye=1.0, T_MeV=10.0, cfl_factor=0.4, flavor_cfl_factor=0.4,
                max_adaptive_speedup=0.0, maxError=1e-6, shape_order=2,
                mass1_eV=0.0, mass2_eV=math.sqrt(7.4e-5), mass3_eV=math.sqrt(2.5e-3),
                theta12_degrees=33.82, theta13_degrees=48.3, theta23_degrees=8.61, deltaCP_degrees=222.0)
    base.update(kw)
    return TestParams.from_dict(base)


def cpu16() -> TestParams:
    """2 flavours, periodic 16^3 cells of 1 cm, 26 directions at 50 MeV (106,496
    particles); isotropic nu, nubar flux factor 1/3 along +z; dm2 = 2.4e-3 eV^2,
    theta = 8.6 deg; 10 RK4 steps at CFL 0.4."""
    return _fast_flavor(nf=2, ncell=(16, 16, 16), Lx=16.0, Ly=16.0, Lz=16.0, directions="cube26",
                        st4_fluxfacbar=1.0 / 3.0, mass1_eV=-math.sqrt(2.4e-3), mass2_eV=0.0, mass3_eV=0.0,
                        theta12_degrees=8.6, theta13_degrees=0.0, theta23_degrees=0.0, deltaCP_degrees=0.0,
                        nsteps=10)


def ff64() -> TestParams:
    """3 flavours, periodic 64^3 cells of 0.5 cm, 92 directions (24,117,248
    particles), fast-flavor setup, 50 RK4 steps: the headline workload."""
    return _fast_flavor(nf=3, ncell=(64, 64, 64), Lx=32.0, Ly=32.0, Lz=32.0, nphi_equator=16, nsteps=50)


def fiducial1d() -> TestParams:
    """3 flavours, 1x1x1024 cells over 64 cm along z, 378 directions (387,072
    particles); single mode k = 2 pi/64 cm^-1 of amplitude 1e-6 on f01 plus
    1e-9 noise; 2000 RK4 steps."""
    return _fast_flavor(nf=3, ncell=(1, 1, 1024), Lx=1.0, Ly=1.0, Lz=64.0, nphi_equator=32, st4_amplitude=1e-9,
                        st3_amplitude=1e-6, st3_wavelength_fraction_of_domain=1, nsteps=2000)


def adjoint32() -> TestParams:
    """2 flavours, periodic 32^3 cells of 0.5 cm, 92 directions (3,014,656
    particles), fast-flavor densities, 200 RK4 steps."""
    return _fast_flavor(nf=2, ncell=(32, 32, 32), Lx=16.0, Ly=16.0, Lz=16.0, nphi_equator=16,
                        mass1_eV=-math.sqrt(7.4e-5), mass2_eV=0.0, mass3_eV=0.0, theta12_degrees=33.82,
                        theta13_degrees=0.0, theta23_degrees=0.0, deltaCP_degrees=0.0, nsteps=200)


def large128() -> TestParams:
    """3 flavours, periodic 128^3 cells of 0.25 cm, 92 directions (192,937,984
    particles), fast-flavor setup, 100 RK4 steps."""
    return _fast_flavor(nf=3, ncell=(128, 128, 128), Lx=32.0, Ly=32.0, Lz=32.0, nphi_equator=16, nsteps=100)


WORKLOADS = {"cpu16": cpu16, "ff64": ff64, "fiducial1d": fiducial1d, "adjoint32": adjoint32, "large128": large128}


def scaled(p: TestParams, ncell: tuple) -> TestParams:
    """The same physics on fewer cells of the same size (for parity tests)."""
    dx = (p.Lx / p.ncell[0], p.Ly / p.ncell[1], p.Lz / p.ncell[2])
    q = TestParams.from_dict({k: getattr(p, k) for k in p.__dataclass_fields__ if k != "extra"})
    q.ncell = tuple(ncell)
    q.Lx, q.Ly, q.Lz = dx[0] * ncell[0], dx[1] * ncell[1], dx[2] * ncell[2]
    return q
