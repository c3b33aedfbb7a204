"""Run parameters: the ParmParse keys of Source/Parameters.H:11-131.

``TestParams`` keeps the deck's units; ``emu_params()`` applies the
reference's conversions (Parameters.H:81-97: masses eV -> g, degrees ->
radians) in the same operation order.  NUM_FLAVORS and SHAPE_FACTOR_ORDER,
compile-time in the reference (Make.Emu:1-2), are fields here.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, fields

from . import capi

# AMReX RungeKuttaTimeIntegrator preset tableaus by integration.rk.type:
# (rows of A below the diagonal, b)
TABLEAUS = {
    1: (((),), (1.0,)),                                                          # forward Euler
    2: (((), (1.0,)), (0.5, 0.5)),                                               # trapezoid
    3: (((), (1.0,), (0.25, 0.25)), (1. / 6., 1. / 6., 2. / 3.)),               # SSPRK3
    4: (((), (0.5,), (0.0, 0.5), (0.0, 0.0, 1.0)), (1. / 6., 1. / 3., 1. / 3., 1. / 6.)),  # RK4
}

EV = 1.60218e-12
C = 2.99792458e10
C2 = C * C


@dataclass
class TestParams:
    nf: int = 2                      # NUM_FLAVORS
    shape_order: int = 2             # SHAPE_FACTOR_ORDER
    simulation_type: int = 4
    ncell: tuple = (1, 1, 1)
    Lx: float = 1.0
    Ly: float = 1.0
    Lz: float = 1.0
    nppc: tuple = (1, 1, 1)
    nphi_equator: int = 1
    directions: str = "sphere"       # sphere = uniform_sphere_xyz(nphi_equator); cube26; fibonacci
    ndirections: int = 0
    max_grid_size: int = 64
    nsteps: int = 1
    end_time: float = 1.0
    rho_g_ccm: float = 0.0
    Ye: float = 1.0
    T_MeV: float = 10.0
    cfl_factor: float = 0.5
    flavor_cfl_factor: float = 0.5
    max_adaptive_speedup: float = 0.0
    maxError: float = 1e-6
    write_plot_every: int = 0
    write_plot_particles_every: int = 0
    do_restart: int = 0
    restart_dir: str = ""
    mass1_eV: float = 0.0
    mass2_eV: float = 0.0
    mass3_eV: float = 0.0
    theta12_degrees: float = 0.0
    theta13_degrees: float = 0.0
    theta23_degrees: float = 0.0
    alpha1_degrees: float = 0.0
    alpha2_degrees: float = 0.0
    deltaCP_degrees: float = 0.0
    st3_amplitude: float = 0.0
    st3_wavelength_fraction_of_domain: int = 1
    st4_ndens: float = 0.0
    st4_theta: float = 0.0
    st4_phi: float = 0.0
    st4_fluxfac: float = 0.0
    st4_ndensbar: float = 0.0
    st4_thetabar: float = 0.0
    st4_phibar: float = 0.0
    st4_fluxfacbar: float = 0.0
    st4_amplitude: float = 0.0
    st5_nnue: float = 0.0
    st5_nnua: float = 0.0
    st5_nnux: float = 0.0
    st5_fxnue: float = 0.0
    st5_fxnua: float = 0.0
    st5_fxnux: float = 0.0
    st5_fynue: float = 0.0
    st5_fynua: float = 0.0
    st5_fynux: float = 0.0
    st5_fznue: float = 0.0
    st5_fznua: float = 0.0
    st5_fznux: float = 0.0
    st5_amplitude: float = 0.0
    seed: int = 1234
    integration_type: int = 1        # integration.type: 0 = ForwardEuler, 1 = RungeKutta
    integration_rk_type: int = 4     # integration.rk.type: 1 Euler, 2 trapezoid, 3 SSPRK3, 4 RK4
    extra: dict = field(default_factory=dict)  # keys this build does not use

    # ---- deck I/O -------------------------------------------------------
    @classmethod
    def from_dict(cls, d: dict, **overrides) -> "TestParams":
        names = {f.name for f in fields(cls)}
        kw, extra = {}, {}
        dotted = {"integration.type": "integration_type", "integration.rk.type": "integration_rk_type"}
        for k, v in {**d, **overrides}.items():
            k = dotted.get(k, k)
            if k in names:
                kw[k] = tuple(v) if isinstance(v, list) else v
            else:
                extra[k] = v
        p = cls(**kw)
        p.extra = extra
        return p

    @classmethod
    def from_file(cls, path: str, **overrides) -> "TestParams":
        """Read an AMReX ParmParse deck (key = value, '#' comments, tuples in
        parentheses), as the reference's main3d...ex inputs_file does."""
        d = {}
        with open(path) as f:
            for line in f:
                line = line.split("#")[0].strip()
                if "=" not in line:
                    continue
                k, v = (s.strip() for s in line.split("=", 1))
                d[k] = _parse_value(v)
        return cls.from_dict(d, **overrides)

    # ---- conversions ----------------------------------------------------
    def mass_g(self, i: int) -> float:
        return (self.mass1_eV, self.mass2_eV, self.mass3_eV)[i] * (EV / C2)

    @staticmethod
    def rad(deg: float) -> float:
        return deg * (math.pi / 180.)

    def emu_params(self) -> capi.EmuParams:
        p = capi.EmuParams()
        p.nf = self.nf
        p.shape_order = self.shape_order
        p.ncell[:] = list(self.ncell)
        p.plo[:] = [0.0, 0.0, 0.0]
        p.phi[:] = [self.Lx, self.Ly, self.Lz]
        p.rho, p.Ye, p.T = self.rho_g_ccm, self.Ye, self.T_MeV
        p.mass[:] = [self.mass_g(0), self.mass_g(1), self.mass_g(2)]
        p.theta12 = self.rad(self.theta12_degrees)
        p.theta13 = self.rad(self.theta13_degrees)
        p.theta23 = self.rad(self.theta23_degrees)
        p.alpha1 = self.rad(self.alpha1_degrees)
        p.alpha2 = self.rad(self.alpha2_degrees)
        p.deltaCP = self.rad(self.deltaCP_degrees)
        p.maxError = self.maxError
        p.cfl_factor = self.cfl_factor
        p.flavor_cfl_factor = self.flavor_cfl_factor
        p.max_adaptive_speedup = self.max_adaptive_speedup
        return p

    def init_config(self) -> capi.EmuInitConfig:
        c = capi.EmuInitConfig()
        c.simulation_type = self.simulation_type
        c.nppc[:] = list(self.nppc)
        c.direction_set = {"sphere": capi.DIRS_SPHERE, "cube26": capi.DIRS_CUBE26,
                           "fibonacci": capi.DIRS_FIBONACCI}[self.directions]
        c.nphi_equator = self.nphi_equator
        c.ndirections = self.ndirections
        c.st3_wavelength_fraction_of_domain = self.st3_wavelength_fraction_of_domain
        c.st3_amplitude = self.st3_amplitude
        for k in ("ndens", "theta", "phi", "fluxfac", "ndensbar", "thetabar", "phibar", "fluxfacbar", "amplitude"):
            setattr(c, "st4_" + k, getattr(self, "st4_" + k))
        st5 = [self.st5_nnue, self.st5_nnua, self.st5_nnux, self.st5_fxnue, self.st5_fxnua, self.st5_fxnux,
               self.st5_fynue, self.st5_fynua, self.st5_fynux, self.st5_fznue, self.st5_fznua, self.st5_fznux,
               self.st5_amplitude]
        c.st5[:] = st5
        c.seed = self.seed
        return c

    def tableau(self):
        """The AMReX RungeKuttaTimeIntegrator preset tableau (A rows, b) the
        deck's integration keys select; integration.type = 0 is forward
        Euler, the one-stage tableau."""
        rk = 1 if self.integration_type == 0 else self.integration_rk_type
        if self.integration_type not in (0, 1) or rk not in TABLEAUS:
            raise ValueError(f"integration.type={self.integration_type}, integration.rk.type="
                             f"{self.integration_rk_type} is not supported")
        return TABLEAUS[rk]

    def st5_dict(self) -> dict:
        return {k[4:]: getattr(self, k) for k in (f.name for f in fields(self)) if k.startswith("st5_")}


def _parse_value(v: str):
    v = v.strip()
    if v.startswith("("):
        return tuple(int(x) for x in v.strip("()").split(","))
    if v.startswith('"'):
        return v.strip('"')
    for conv in (int, float):
        try:
            return conv(v)
        except ValueError:
            pass
    return v
