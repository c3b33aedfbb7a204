"""Python face of the FP64 CPU oracle.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module; the product path
(``emu_amd``) never does.

Contents
  * ctypes bindings of ``liboracle.so`` (``emu_oracle.c``): the AoS/ghost-cell
    restatement of Source/Evolve.cpp, FlavoredNeutrinoContainer.cpp and the
    AMReX TimeIntegrator RK4 step.
  * ``init_particles``: a numpy restatement of
    Source/FlavoredNeutrinoContainerInit.cpp:136-661 (simulation types 0-5,
    ``uniform_sphere_xyz``, Minerbo closure, ``set_trace_length``).  The
    reference draws its perturbations from amrex::Random seeded with the wall
    clock (Source/main.cpp:238-240); here every draw comes from Philox4x32-10
    (Salmon et al., SC'11) keyed by (seed, particle id, draw index), so runs are
    reproducible and the GPU path can be fed identical arrays.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

# Constants.H:6-26, same operation order as the constexprs.
eV = 1.60218e-12
clight = 2.99792458e10
hbar = 1.05457266e-27
c2 = clight * clight
c4 = c2 * c2
hbarc = hbar * clight
GF = 1.1663787e-5 / (1e9 * 1e9 * eV * eV) * hbarc * hbarc * hbarc
Mp = 1.6726219e-24

OK, ERR_TRACE, ERR_NEGDIAG, ERR_LENGTH, ERR_DT = 0, 1, 2, 3, 4


def nattribs(nf: int) -> int:
    return 12 + 2 * nf * nf


def ncomp(nf: int) -> int:
    return 3 + 8 * nf * nf


def herm_names(nf: int, template: str) -> list[str]:
    """Component names in HermitianUtils.py:77-83,170-177 order."""
    out = []
    for i in range(nf):
        for j in range(i, nf):
            out.append(template.format(i, j, "Re"))
            if j > i:
                out.append(template.format(i, j, "Im"))
    return out


def attribute_names(nf: int) -> list[str]:
    """PIdx order (FlavoredNeutrinoContainer.H:22-24, generate_code.py:93-104)."""
    names = ["time", "x", "y", "z", "pupx", "pupy", "pupz", "pupt"]
    for t in ("", "bar"):
        names += ["N" + t, "L" + t] + herm_names(nf, "f{}{}_{}" + t)
    return names


def grid_names(nf: int) -> list[str]:
    """GIdx order (Evolve.H:18-22, generate_code.py:127-135)."""
    names = ["rho", "T", "Ye"]
    for v in ("N", "Fx", "Fy", "Fz"):
        for t in ("", "bar"):
            names += herm_names(nf, v + "{}{}_{}" + t)
    return names


def particle_dtype(nf: int) -> np.dtype:
    return np.dtype([("pos", "<f8", (3,)), ("rdata", "<f8", (nattribs(nf),)),
                     ("id", "<i4"), ("cpu", "<i4")])


class OrParams(ctypes.Structure):
    _fields_ = [("ncell", ctypes.c_int * 3), ("shape_order", ctypes.c_int),
                ("plo", ctypes.c_double * 3), ("phi", ctypes.c_double * 3),
                ("rho", ctypes.c_double), ("Ye", ctypes.c_double), ("T", ctypes.c_double),
                ("mass", ctypes.c_double * 3),
                ("theta12", ctypes.c_double), ("theta13", ctypes.c_double),
                ("theta23", ctypes.c_double), ("alpha1", ctypes.c_double),
                ("alpha2", ctypes.c_double), ("deltaCP", ctypes.c_double),
                ("maxError", ctypes.c_double), ("cfl_factor", ctypes.c_double),
                ("flavor_cfl_factor", ctypes.c_double),
                ("max_adaptive_speedup", ctypes.c_double), ("nthreads", ctypes.c_int)]


@dataclass
class Deck:
    """The ParmParse keys of Source/Parameters.H:52-131, in deck units."""
    nf: int = 2
    shape_order: int = 2
    simulation_type: int = 4
    ncell: tuple = (1, 1, 1)
    Lx: float = 1.0
    Ly: float = 1.0
    Lz: float = 1.0
    nppc: tuple = (1, 1, 1)
    nphi_equator: int = 1
    directions: str = "sphere"      # "sphere" (uniform_sphere_xyz), "cube26", "fibonacci"
    ndirections: int = 0            # for "fibonacci"
    nsteps: int = 1
    end_time: float = 1.0
    rho_g_ccm: float = 0.0
    Ye: float = 1.0
    T_MeV: float = 10.0
    cfl_factor: float = 0.5
    flavor_cfl_factor: float = 0.5
    max_adaptive_speedup: float = 0.0
    maxError: float = 1e-6
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
    st5: dict = field(default_factory=dict)
    seed: int = 1234

    # Parameters.H:81-97 unit conversions, same operation order.
    def mass_g(self, i: int) -> float:
        m = (self.mass1_eV, self.mass2_eV, self.mass3_eV)[i]
        return m * (eV / c2)

    def rad(self, deg: float) -> float:
        return deg * (math.pi / 180.)

    def params(self, nthreads: int = 1) -> OrParams:
        p = OrParams()
        p.ncell[:] = list(self.ncell)
        p.shape_order = self.shape_order
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
        p.nthreads = nthreads
        return p


# ---------------------------------------------------------------------------
# Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), vectorised over counters.
# ---------------------------------------------------------------------------
_PH_M0, _PH_M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_PH_W0, _PH_W1 = 0x9E3779B9, 0xBB67AE85
_M32 = np.uint64(0xFFFFFFFF)


def philox4x32(c0, c1, c2_, c3, k0: int, k1: int):
    c0 = np.asarray(c0, dtype=np.uint64) & _M32
    c1 = np.asarray(c1, dtype=np.uint64) & _M32
    c2_ = np.asarray(c2_, dtype=np.uint64) & _M32
    c3 = np.asarray(c3, dtype=np.uint64) & _M32
    k0, k1 = k0 & 0xFFFFFFFF, k1 & 0xFFFFFFFF
    for _ in range(10):
        p0 = _PH_M0 * c0
        p1 = _PH_M1 * c2_
        hi0, lo0 = p0 >> np.uint64(32), p0 & _M32
        hi1, lo1 = p1 >> np.uint64(32), p1 & _M32
        c0, c1, c2_, c3 = (hi1 ^ c1 ^ np.uint64(k0)), lo1, (hi0 ^ c3 ^ np.uint64(k1)), lo0
        k0 = (k0 + _PH_W0) & 0xFFFFFFFF
        k1 = (k1 + _PH_W1) & 0xFFFFFFFF
    return c0, c1, c2_, c3


def uniform01(seed: int, pid, draw: int):
    """Uniform [0,1) double for (seed, particle id, draw index): 53 bits from
    two 32-bit Philox outputs of counter (pid_lo, pid_hi, draw>>1, 0)."""
    pid = np.asarray(pid, dtype=np.uint64)
    r = philox4x32(pid & _M32, pid >> np.uint64(32), np.full_like(pid, draw >> 1),
                   np.zeros_like(pid), seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    hi, lo = (r[0], r[1]) if (draw & 1) == 0 else (r[2], r[3])
    return ((hi >> np.uint64(5)).astype(np.float64) * 67108864.0
            + (lo >> np.uint64(6)).astype(np.float64)) * (1.0 / 9007199254740992.0)


def symmetric_uniform(seed, pid, draw):
    """FlavoredNeutrinoContainerInit.cpp:111-113."""
    return 2. * (uniform01(seed, pid, draw) - 0.5)


# ---------------------------------------------------------------------------
# Direction sets
# ---------------------------------------------------------------------------
def uniform_sphere_xyz(nphi_at_equator: int) -> np.ndarray:
    """FlavoredNeutrinoContainerInit.cpp:9-37."""
    dtheta = math.pi * math.sqrt(3) / nphi_at_equator
    xyz = []
    theta = 0.0
    phi0 = 0.0
    while theta < math.pi / 2.:
        nphi = nphi_at_equator if theta == 0 else int(math.floor(nphi_at_equator * math.cos(theta) + 0.5))
        dphi = 2. * math.pi / nphi
        if nphi == 1:
            theta = math.pi / 2.
        for iphi in range(nphi):
            phi = phi0 + iphi * dphi
            x = math.cos(theta) * math.cos(phi)
            y = math.cos(theta) * math.sin(phi)
            z = math.sin(theta)
            xyz.append((x, y, z))
            if theta > 0:
                xyz.append((-x, -y, -z))
        theta += dtheta
        phi0 = phi0 + 0.5 * dphi
    return np.array(xyz, dtype=np.float64)


def cube26() -> np.ndarray:
    """6 face + 12 edge + 8 corner unit vectors of a cube (octahedral+diagonal set)."""
    out = []
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if dx == dy == dz == 0:
                    continue
                n = math.sqrt(dx * dx + dy * dy + dz * dz)
                out.append((dx / n, dy / n, dz / n))
    return np.array(out, dtype=np.float64)


def fibonacci_sphere(n: int) -> np.ndarray:
    """Fibonacci lattice on the sphere (golden-angle spiral)."""
    out = []
    ga = math.pi * (3.0 - math.sqrt(5.0))
    for i in range(n):
        z = 1.0 - (2.0 * i + 1.0) / n
        r = math.sqrt(max(0.0, 1.0 - z * z))
        out.append((r * math.cos(ga * i), r * math.sin(ga * i), z))
    return np.array(out, dtype=np.float64)


def direction_set(deck: Deck) -> np.ndarray:
    if deck.directions == "sphere":
        return uniform_sphere_xyz(deck.nphi_equator)
    if deck.directions == "cube26":
        return cube26()
    if deck.directions == "fibonacci":
        return fibonacci_sphere(deck.ndirections)
    raise ValueError(deck.directions)


def su_magnitude(f: np.ndarray, nf: int) -> np.ndarray:
    """HermitianUtils.py:102-123 on the last axis (numpy restatement)."""
    names = herm_names(nf, "{}{}{}")
    idx = {n: k for k, n in enumerate(names)}
    m2 = np.zeros(f.shape[:-1])
    for i in range(nf):
        for j in range(i + 1, nf):
            m2 = m2 + f[..., idx[f"{i}{j}Re"]] ** 2 + f[..., idx[f"{i}{j}Im"]] ** 2
    for l in range(1, nf):
        bc = np.zeros(f.shape[:-1])
        for i in range(1, l + 1):
            bc = bc + f[..., idx[f"{i-1}{i-1}Re"]]
        bc = bc - l * f[..., idx[f"{l}{l}Re"]]
        bc = bc * math.sqrt(2.0 / (l * (l + 1.0)))
        m2 = m2 + (bc / 2.0) ** 2
    return np.sqrt(m2)


def minerbo_Z(fluxfac: float) -> float:
    """FlavoredNeutrinoContainerInit.cpp:46-80."""
    maxresidual, maxcount, minfluxfac = 1e-6, 20, 1e-3
    Z = 1.0
    if fluxfac < minfluxfac:
        return 3. * fluxfac
    residual, count = 1.0, 0
    while abs(residual) > maxresidual and count < maxcount:
        residual = fluxfac - 1.0 / math.tanh(Z) + 1.0 / Z
        slope = 1.0 / (math.sinh(Z) * math.sinh(Z)) - 1.0 / (Z * Z)
        Z -= residual / slope
        count += 1
    if residual > maxresidual:
        raise RuntimeError("Failed to converge on a solution.")
    return Z


def init_particles(deck: Deck) -> np.ndarray:
    """InitParticles, FlavoredNeutrinoContainerInit.cpp:136-655, one tile
    covering the domain: particles ordered by cell (z fastest, :222), then
    location in the cell, then direction (:294)."""
    nf = deck.nf
    dtype = particle_dtype(nf)
    names = attribute_names(nf)
    K = {n: i for i, n in enumerate(names)}
    nx, ny, nz = deck.ncell
    L = (deck.Lx, deck.Ly, deck.Lz)
    dx = [L[d] / float(deck.ncell[d]) for d in range(3)]
    dirs = direction_set(deck)
    ndirs = len(dirs)
    nlocs = deck.nppc[0] * deck.nppc[1] * deck.nppc[2]
    scale_fac = dx[0] * dx[1] * dx[2] / nlocs / ndirs  # :156

    # cell-major (ix, iy, iz), z fastest; then location; then direction
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
    npx, npy, npz = deck.nppc
    locs = []
    for i_part in range(nlocs):  # get_position_unit_cell, :84-97
        ixp = i_part // (npy * npz)
        iyp = (i_part % (npy * npz)) % npy
        izp = (i_part % (npy * npz)) // npy
        locs.append(((0.5 + ixp) / npx, (0.5 + iyp) / npy, (0.5 + izp) / npz))
    ncells = nx * ny * nz
    npart = ncells * nlocs * ndirs
    p = np.zeros(npart, dtype=dtype)
    cell_of = np.repeat(np.arange(ncells), nlocs * ndirs)
    loc_of = np.tile(np.repeat(np.arange(nlocs), ndirs), ncells)
    dir_of = np.tile(np.arange(ndirs), ncells * nlocs)
    r = np.array(locs)[loc_of]
    x = 0.0 + (ix[cell_of] + r[:, 0]) * dx[0]
    y = 0.0 + (iy[cell_of] + r[:, 1]) * dx[1]
    z = 0.0 + (iz[cell_of] + r[:, 2]) * dx[2]
    pid = np.arange(npart, dtype=np.int64)
    p["id"] = (pid + 1).astype(np.int32)
    p["cpu"] = 0
    p["pos"][:, 0], p["pos"][:, 1], p["pos"][:, 2] = x, y, z
    rd = p["rdata"]
    rd[:, K["x"]], rd[:, K["y"]], rd[:, K["z"]] = x, y, z
    rd[:, K["time"]] = 0
    u = dirs[dir_of]
    st = deck.simulation_type
    m1, m2_, m3 = deck.mass_g(0), deck.mass_g(1), deck.mass_g(2)
    th12 = deck.rad(deck.theta12_degrees)
    dm2 = (m2_ - m1) * (m2_ - m1)
    seed = deck.seed

    def setf(name, val, bar=False):
        rd[:, K[name + ("bar" if bar else "")]] = val

    if st == 0:
        rd[:, K["N"]] = 1.0
        rd[:, K["Nbar"]] = 1.0
        setf("f00_Re", 1.0)
        setf("f00_Re", 1.0, True)
        pupt = dm2 * c4 * math.sin(2. * th12) / (8. * math.pi * hbarc)
        rd[:, K["pupt"]] = pupt
    elif st in (1, 2, 3, 4):
        setf("f00_Re", 1.0)
        setf("f00_Re", 1.0, True)
        rd[:, K["pupt"]] = 50. * 1e6 * eV
    if st in (0, 1, 2, 3, 4, 5):
        if st == 5:
            rd[:, K["pupt"]] = 50. * 1e6 * eV
        rd[:, K["pupx"]] = u[:, 0] * rd[:, K["pupt"]]
        rd[:, K["pupy"]] = u[:, 1] * rd[:, K["pupt"]]
        rd[:, K["pupz"]] = u[:, 2] * rd[:, K["pupt"]]
    else:
        raise ValueError("Invalid simulation type")
    pupt = rd[:, K["pupt"]]
    if st == 1:
        ndens = 10. * dm2 * c4 / (2. * math.sqrt(2.) * GF * pupt)
        rd[:, K["N"]] = ndens * scale_fac
        rd[:, K["Nbar"]] = ndens * scale_fac
    elif st == 2:
        omega = dm2 * c4 / (2. * pupt)
        mu_ndens = math.sqrt(2.) * GF
        ndens = omega / (2. * mu_ndens)
        rd[:, K["N"]] = ndens * scale_fac * (1. + u[:, 2])
        rd[:, K["Nbar"]] = ndens * scale_fac * (1. - u[:, 2])
    elif st == 3:
        lam = L[2] / float(deck.st3_wavelength_fraction_of_domain)
        k = (2. * math.pi) / lam
        val = deck.st3_amplitude * np.sin(k * p["pos"][:, 2])
        setf("f01_Re", val)
        setf("f01_Re", val, True)
        omega = dm2 * c4 / (2. * pupt)
        mu_ndens = math.sqrt(2.) * GF
        ndens = (omega + k * hbarc) / (2. * mu_ndens)
        rd[:, K["N"]] = ndens * scale_fac * (1. + u[:, 2])
        rd[:, K["Nbar"]] = ndens * scale_fac * (1. - u[:, 2])
    elif st == 4:
        a = deck.st4_amplitude
        r1, r2, r3, r4 = (symmetric_uniform(seed, pid, d) for d in range(4))
        setf("f01_Re", a * r1)
        setf("f01_Im", a * r2)
        setf("f01_Re", a * r3, True)
        setf("f01_Im", a * r4, True)
        if nf == 3:
            r1, r2, r3, r4 = (symmetric_uniform(seed, pid, d) for d in range(4, 8))
            setf("f02_Re", a * r1)
            setf("f02_Im", a * r2)
            setf("f02_Re", a * r3, True)
            setf("f02_Im", a * r4, True)
        if deck.st3_amplitude != 0.0:  # single mode on top of the noise
            k = (2. * math.pi) / (L[2] / float(deck.st3_wavelength_fraction_of_domain))
            v = deck.st3_amplitude * np.sin(k * p["pos"][:, 2])
            rd[:, K["f01_Re"]] += v
            rd[:, K["f01_Rebar"]] += v
        fh = np.array([math.cos(deck.st4_phi) * math.sin(deck.st4_theta),
                       math.sin(deck.st4_phi) * math.sin(deck.st4_theta), math.cos(deck.st4_theta)])
        fhb = np.array([math.cos(deck.st4_phibar) * math.sin(deck.st4_thetabar),
                        math.sin(deck.st4_phibar) * math.sin(deck.st4_thetabar), math.cos(deck.st4_thetabar)])
        cost = fh[0] * u[:, 0] + fh[1] * u[:, 1] + fh[2] * u[:, 2]
        costb = fhb[0] * u[:, 0] + fhb[1] * u[:, 1] + fhb[2] * u[:, 2]
        rd[:, K["N"]] = deck.st4_ndens * scale_fac * (1. + 3. * deck.st4_fluxfac * cost)
        rd[:, K["Nbar"]] = deck.st4_ndensbar * scale_fac * (1. + 3. * deck.st4_fluxfacbar * costb)
    elif st == 5:
        s5 = deck.st5
        fe = math.sqrt(s5["fxnue"] ** 2 + s5["fynue"] ** 2 + s5["fznue"] ** 2)
        fa = math.sqrt(s5["fxnua"] ** 2 + s5["fynua"] ** 2 + s5["fznua"] ** 2)
        fx = math.sqrt(s5["fxnux"] ** 2 + s5["fynux"] ** 2 + s5["fznux"] ** 2)
        Ze, Za, Zx = minerbo_Z(fe), minerbo_Z(fa), minerbo_Z(fx)

        def mu(fxk, fyk, fzk, ff):
            if ff > 0:
                return (s5[fxk] * u[:, 0] + s5[fyk] * u[:, 1] + s5[fzk] * u[:, 2]) / ff
            return np.zeros(len(u))

        def closure(Z, m):
            res = np.exp(Z * m)
            if Z / 3.0 > 1e-3:
                res = res * (Z / math.sinh(Z))
            return res

        Ne = s5["nnue"] * scale_fac * closure(Ze, mu("fxnue", "fynue", "fznue", fe))
        Na = s5["nnua"] * scale_fac * closure(Za, mu("fxnua", "fynua", "fznua", fa))
        Nx = s5["nnux"] * scale_fac * closure(Zx, mu("fxnux", "fynux", "fznux", fx)) / 4.0
        N = Ne + Nx
        Nb = Na + Nx
        if nf == 3:
            N = N + Nx
            Nb = Nb + Nx
        rd[:, K["N"]], rd[:, K["Nbar"]] = N, Nb
        setf("f00_Re", Ne / N)
        setf("f11_Re", Nx / N)
        setf("f00_Re", Na / Nb, True)
        setf("f11_Re", Nx / Nb, True)
        if nf == 3:
            setf("f22_Re", Nx / N)
            setf("f22_Re", Nx / Nb, True)
        amp = s5["amplitude"]
        d = 0
        pairs = [("01", "00", "11")] + ([("02", "00", "22"), ("12", "11", "22")] if nf == 3 else [])
        order = []
        for (ij, a_, b_) in pairs[:1]:
            order += [(ij, "Re", False, a_, b_), (ij, "Im", False, a_, b_),
                      (ij, "Re", True, a_, b_), (ij, "Im", True, a_, b_)]
        if nf == 3:
            for (ij, a_, b_) in pairs[1:]:
                order += [(ij, "Re", False, a_, b_), (ij, "Im", False, a_, b_)]
            for (ij, a_, b_) in pairs[1:]:
                order += [(ij, "Re", True, a_, b_), (ij, "Im", True, a_, b_)]
        for (ij, part, bar, a_, b_) in order:
            t = "bar" if bar else ""
            rnd = symmetric_uniform(seed, pid, d)
            d += 1
            rd[:, K[f"f{ij}_{part}{t}"]] = amp * rnd * (rd[:, K[f"f{a_}_Re{t}"]] - rd[:, K[f"f{b_}_Re{t}"]])
    # set_trace_length, generate_code.py:450-457
    fn = herm_names(nf, "f{}{}_{}")
    rd[:, K["L"]] = su_magnitude(rd[:, [K[n] for n in fn]], nf)
    rd[:, K["Lbar"]] = su_magnitude(rd[:, [K[n + "bar"] for n in fn]], nf)
    return p


# ---------------------------------------------------------------------------
# ctypes bindings
# ---------------------------------------------------------------------------
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `make -C oracle`")
        L_ = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER(OrParams)
        vp, dp, i64 = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int64
        for nf in ("nf2", "nf3"):
            sig = {
                "or_grid_ndoubles": (i64, [P]),
                "or_mass_matrix": (None, [P, dp]),
                "or_su_magnitude": (ctypes.c_double, [dp]),
                "or_commutator_rhs": (None, [dp, dp, dp]),
                "or_grid_init": (None, [P, dp]),
                "or_grid_dims": (None, [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
                "or_deposit_to_mesh": (None, [P, vp, i64, dp]),
                "or_fill_boundary": (None, [P, dp]),
                "or_interpolate_rhs_from_mesh": (None, [P, vp, i64, dp]),
                "or_compute_dt": (ctypes.c_int, [P, dp, ctypes.c_double, dp]),
                "or_vvac_max": (ctypes.c_double, [P, vp, i64]),
                "or_renormalize": (ctypes.c_int, [P, vp, i64, ctypes.POINTER(i64)]),
                "or_sync_location": (None, [vp, i64, ctypes.c_int]),
                "or_redistribute": (None, [P, vp, i64]),
                "or_apply_rhs": (None, [vp, i64, ctypes.c_double, vp]),
                "or_rk4_step": (ctypes.c_int, [P, vp, i64, dp, ctypes.c_double, ctypes.c_double, dp,
                                               ctypes.POINTER(i64)]),
                "or_rk_step": (ctypes.c_int, [P, vp, i64, dp, ctypes.c_double, ctypes.c_double, ctypes.c_int, dp,
                                              dp, dp, ctypes.POINTER(i64)]),
                "or_evolve": (ctypes.c_int, [P, vp, i64, dp, ctypes.c_int, ctypes.c_double, dp, dp, dp]),
            }
            for name, (res, args) in sig.items():
                fn = getattr(L_, f"{name}_{nf}")
                fn.restype, fn.argtypes = res, args
        L_.or_interp.restype = None
        L_.or_interp.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int), dp]
        L_.or_constants.restype = None
        L_.or_constants.argtypes = [dp]
        L_.or_roundoff_hi.restype = None
        L_.or_roundoff_hi.argtypes = [P, dp]
        _lib = L_
    return _lib


def _f(nf, name):
    return getattr(lib(), f"{name}_nf{nf}")


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def interp(delta: float, order: int):
    first = ctypes.c_int()
    w = np.zeros(3)
    lib().or_interp(delta, order, ctypes.byref(first), _dp(w))
    return first.value, w


def constants() -> np.ndarray:
    out = np.zeros(5)
    lib().or_constants(_dp(out))
    return out


def roundoff_hi(params: OrParams) -> np.ndarray:
    out = np.zeros(3)
    lib().or_roundoff_hi(ctypes.byref(params), _dp(out))
    return out


def mass_matrix(nf: int, params: OrParams) -> np.ndarray:
    out = np.zeros(nf * nf)
    _f(nf, "or_mass_matrix")(ctypes.byref(params), _dp(out))
    return out


def su_mag(nf: int, h) -> float:
    h = np.ascontiguousarray(h, dtype=np.float64)
    return _f(nf, "or_su_magnitude")(_dp(h))


def commutator_rhs(nf: int, H, f) -> np.ndarray:
    H = np.ascontiguousarray(H, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros(nf * nf)
    _f(nf, "or_commutator_rhs")(_dp(H), _dp(f), _dp(out))
    return out


class Grid:
    """The reference's MultiFab `state`: one periodic fab with ghost cells,
    component-major [comp][k][j][i] (main.cpp:67-85)."""

    def __init__(self, nf: int, params: OrParams):
        self.nf = nf
        self.params = params
        n = _f(nf, "or_grid_ndoubles")(ctypes.byref(params))
        self.data = np.zeros(n, dtype=np.float64)
        ng = (ctypes.c_int * 3)()
        nfab = (ctypes.c_int * 3)()
        _f(nf, "or_grid_dims")(ctypes.byref(params), ng, nfab)
        self.ng = tuple(ng)
        self.nfab = tuple(nfab)
        _f(nf, "or_grid_init")(ctypes.byref(params), _dp(self.data))

    def fab(self) -> np.ndarray:
        """[comp, k, j, i] view including ghost cells."""
        return self.data.reshape(ncomp(self.nf), self.nfab[2], self.nfab[1], self.nfab[0])

    def valid(self) -> np.ndarray:
        """[comp, k, j, i] view of the valid cells."""
        g = self.ng
        n = self.params.ncell
        return self.fab()[:, g[2]:g[2] + n[2], g[1]:g[1] + n[1], g[0]:g[0] + n[0]]


def deposit_to_mesh(nf, params, parts, grid: Grid):
    _f(nf, "or_deposit_to_mesh")(ctypes.byref(params), parts.ctypes.data, len(parts), _dp(grid.data))


def fill_boundary(nf, params, grid: Grid):
    _f(nf, "or_fill_boundary")(ctypes.byref(params), _dp(grid.data))


def interpolate_rhs_from_mesh(nf, params, rhs, grid: Grid):
    _f(nf, "or_interpolate_rhs_from_mesh")(ctypes.byref(params), rhs.ctypes.data, len(rhs), _dp(grid.data))


def compute_dt(nf, params, grid: Grid, vvac: float):
    dt = ctypes.c_double()
    err = _f(nf, "or_compute_dt")(ctypes.byref(params), _dp(grid.data), vvac, ctypes.byref(dt))
    return err, dt.value


def vvac_max(nf, params, parts) -> float:
    return _f(nf, "or_vvac_max")(ctypes.byref(params), parts.ctypes.data, len(parts))


def renormalize(nf, params, parts):
    bad = ctypes.c_int64(-1)
    err = _f(nf, "or_renormalize")(ctypes.byref(params), parts.ctypes.data, len(parts), ctypes.byref(bad))
    return err, bad.value


def sync_location(nf, parts, type_):
    _f(nf, "or_sync_location")(parts.ctypes.data, len(parts), type_)


def redistribute(nf, params, parts):
    _f(nf, "or_redistribute")(ctypes.byref(params), parts.ctypes.data, len(parts))


def apply_rhs(nf, parts, a, dF):
    _f(nf, "or_apply_rhs")(parts.ctypes.data, len(parts), a, dF.ctypes.data)


def rk4_step(nf, params, parts, grid: Grid, dt: float, vvac: float):
    dtn = ctypes.c_double()
    bad = ctypes.c_int64(-1)
    err = _f(nf, "or_rk4_step")(ctypes.byref(params), parts.ctypes.data, len(parts), _dp(grid.data), dt,
                                vvac, ctypes.byref(dtn), ctypes.byref(bad))
    return err, dtn.value, bad.value


# AMReX RungeKuttaTimeIntegrator preset tableaus by integration.rk.type (A row-major, b);
# integration.type = 0 (the ForwardEuler integrator) is the 1-stage tableau.
TABLEAUS = {
    1: ([[0.0]], [1.0]),                                                     # forward Euler
    2: ([[0.0, 0.0], [1.0, 0.0]], [0.5, 0.5]),                               # trapezoid
    3: ([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.25, 0.25, 0.0]], [1. / 6., 1. / 6., 2. / 3.]),  # SSPRK3
    4: ([[0.0, 0.0, 0.0, 0.0], [0.5, 0.0, 0.0, 0.0], [0.0, 0.5, 0.0, 0.0], [0.0, 0.0, 1.0, 0.0]],
        [1. / 6., 1. / 3., 1. / 3., 1. / 6.]),                                # RK4
}


def rk_step(nf, params, parts, grid: Grid, dt: float, vvac: float, rk_type: int):
    A, b = TABLEAUS[rk_type]
    n = len(b)
    Aa = np.ascontiguousarray(np.array(A, dtype=np.float64).reshape(n * n))
    ba = np.ascontiguousarray(np.array(b, dtype=np.float64))
    dtn = ctypes.c_double()
    bad = ctypes.c_int64(-1)
    err = _f(nf, "or_rk_step")(ctypes.byref(params), parts.ctypes.data, len(parts), _dp(grid.data), dt, vvac, n,
                               _dp(Aa), _dp(ba), ctypes.byref(dtn), ctypes.byref(bad))
    return err, dtn.value, bad.value


def evolve(nf, params, parts, grid: Grid, nsteps: int, end_time: float = 1e300):
    t = ctypes.c_double(0.0)
    dt = ctypes.c_double(0.0)
    log = np.zeros(max(nsteps, 1))
    err = _f(nf, "or_evolve")(ctypes.byref(params), parts.ctypes.data, len(parts), _dp(grid.data), nsteps,
                              end_time, ctypes.byref(t), ctypes.byref(dt), _dp(log))
    return err, t.value, dt.value, log
