"""Python face of the reference's call surface over the C ABI.

Mirrors Source/FlavoredNeutrinoContainer.H, Evolve.H and main.cpp:
``FlavoredNeutrinoContainer`` (SoA device arrays in PIdx order),
``MultiFab`` (the grid `state` as one ghost-cell fab in GIdx order) and the
free functions ``deposit_to_mesh``, ``interpolate_rhs_from_mesh``,
``compute_dt``.  ``FusedState`` is the library-owned, cell-sorted state that
advances whole RK4 steps.  torch provides the device memory and the stream;
every computation happens in libemu_amd.so.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import capi
from .capi import check, lib
from .params import TestParams


def attribute_names(nf: int) -> list:
    """PIdx order (FlavoredNeutrinoContainer.H:12-26, generate_code.py:93-104)."""
    def herm(t):
        out = []
        for i in range(nf):
            for j in range(i, nf):
                out.append(t.format(i, j, "Re"))
                if j > i:
                    out.append(t.format(i, j, "Im"))
        return out
    names = ["time", "x", "y", "z", "pupx", "pupy", "pupz", "pupt"]
    for t in ("", "bar"):
        names += ["N" + t, "L" + t] + herm("f{}{}_{}" + t)
    return names


class Emu:
    """An emu_ctx: geometry, mass matrix and scratch for one parameter set,
    bound to torch's current stream on the current device."""

    def __init__(self, parms: TestParams, device: str = "cuda"):
        if not torch.cuda.is_available():
            raise RuntimeError("emu_amd needs a GPU (gfx950); none is visible")
        self.parms = parms
        self.nf = parms.nf
        self.nh = parms.nf * parms.nf
        self.nattribs = 12 + 2 * self.nh
        self.ncomp = 3 + 8 * self.nh
        self.device = torch.device(device)
        self._params = parms.emu_params()
        ctx = ctypes.c_void_p()
        check(lib().emu_create(ctypes.byref(self._params), ctypes.byref(ctx)), "emu_create")
        self.ctx = ctx
        self.bind_stream()
        self.names = attribute_names(self.nf)
        self.K = {n: i for i, n in enumerate(self.names)}

    def bind_stream(self, stream: torch.cuda.Stream | None = None):
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        check(lib().emu_set_stream(self.ctx, ctypes.c_void_p(s.cuda_stream)), "emu_set_stream", self.ctx)

    def sync(self):
        check(lib().emu_sync(self.ctx), "emu_sync", self.ctx)

    def close(self):
        if getattr(self, "ctx", None):
            lib().emu_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def ngrow(self):
        g = (ctypes.c_int * 3)()
        check(lib().emu_grid_ngrow(self.ctx, g), "emu_grid_ngrow", self.ctx)
        return tuple(g)


class FlavoredNeutrinoContainer:
    """SoA particle container on the device: pos (3, np) and rdata
    (nattribs, np), float64, PIdx order; ids (np,) int64."""

    def __init__(self, emu: Emu, np_: int):
        self.emu = emu
        self.np = int(np_)
        dev = emu.device
        self.pos = torch.zeros((3, self.np), dtype=torch.float64, device=dev)
        self.rdata = torch.zeros((emu.nattribs, self.np), dtype=torch.float64, device=dev)
        self.ids = torch.zeros(self.np, dtype=torch.int64, device=dev)

    def view(self) -> capi.EmuParticles:
        v = capi.EmuParticles()
        v.np = self.np
        for d in range(3):
            v.pos[d] = self.pos[d].data_ptr() if self.np else None
        for k in range(self.emu.nattribs):
            v.rdata[k] = self.rdata[k].data_ptr() if self.np else None
        return v

    @classmethod
    def init_particles(cls, emu: Emu) -> "FlavoredNeutrinoContainer":
        """FlavoredNeutrinoContainer::InitParticles (FlavoredNeutrinoContainerInit.cpp:136-661)."""
        cfg = emu.parms.init_config()
        n = lib().emu_init_count(ctypes.byref(emu._params), ctypes.byref(cfg))
        c = cls(emu, n)
        v = c.view()
        check(lib().emu_init_particles(emu.ctx, ctypes.byref(cfg), ctypes.byref(v),
                                       ctypes.c_void_p(c.ids.data_ptr() if n else None)), "InitParticles", emu.ctx)
        return c

    def SyncLocation(self, type_: int):
        v = self.view()
        check(lib().emu_sync_location(self.emu.ctx, ctypes.byref(v), type_), "SyncLocation", self.emu.ctx)

    def RedistributeLocal(self):
        v = self.view()
        check(lib().emu_redistribute_local(self.emu.ctx, ctypes.byref(v)), "RedistributeLocal", self.emu.ctx)

    def Renormalize(self):
        v = self.view()
        bad = ctypes.c_int64(-1)
        code = lib().emu_renormalize(self.emu.ctx, ctypes.byref(v), ctypes.byref(bad))
        check(code, "Renormalize", self.emu.ctx, bad.value)

    def copyParticles(self, other: "FlavoredNeutrinoContainer"):
        if other.np != self.np:
            raise ValueError("size mismatch")
        d, s = self.view(), other.view()
        check(lib().emu_copy_particles(self.emu.ctx, ctypes.byref(d), ctypes.byref(s)), "copyParticles", self.emu.ctx)
        self.ids.copy_(other.ids)

    def apply_rhs(self, a: float, dFdt: "FlavoredNeutrinoContainer"):
        """ApplyFlavoredNeutrinoRHS (FlavoredNeutrinoContainer.H:33-47)."""
        v, d = self.view(), dFdt.view()
        check(lib().emu_apply_rhs(self.emu.ctx, ctypes.byref(v), a, ctypes.byref(d)), "apply_rhs", self.emu.ctx)

    def rk_combine(self, src: "FlavoredNeutrinoContainer", coeffs, Fs):
        """self = src, then self = a_j F_j + self for each (a_j, F_j), in one
        pass: the stage combination of the RK integrator (emu_rk_combine)."""
        coeffs = [float(a) for a in coeffs]
        if len(coeffs) != len(Fs) or any(F.np != self.np for F in Fs) or src.np != self.np:
            raise ValueError("size mismatch")
        n = len(coeffs)
        a = (ctypes.c_double * max(n, 1))(*coeffs)
        views = (capi.EmuParticles * max(n, 1))(*[F.view() for F in Fs])
        d, sv = self.view(), src.view()
        check(lib().emu_rk_combine(self.emu.ctx, ctypes.byref(d), ctypes.byref(sv), n, a, views), "rk_combine",
              self.emu.ctx)
        self.ids.copy_(src.ids)

    def Vvac_max(self) -> float:
        v = self.view()
        out = ctypes.c_double()
        check(lib().emu_vvac_max(self.emu.ctx, ctypes.byref(v), ctypes.byref(out)), "Vvac_max", self.emu.ctx)
        return out.value

    # ---- host transfer (AoS, the reference's particle struct) -------------
    def to_aos(self) -> np.ndarray:
        dt = np.dtype([("pos", "<f8", (3,)), ("rdata", "<f8", (self.emu.nattribs,)), ("id", "<i4"), ("cpu", "<i4")])
        out = np.zeros(self.np, dtype=dt)
        out["pos"] = self.pos.t().cpu().numpy()
        out["rdata"] = self.rdata.t().cpu().numpy()
        out["id"] = self.ids.cpu().numpy().astype(np.int32)
        return out

    @classmethod
    def from_aos(cls, emu: Emu, arr: np.ndarray) -> "FlavoredNeutrinoContainer":
        c = cls(emu, len(arr))
        if len(arr):
            c.pos.copy_(torch.from_numpy(np.ascontiguousarray(arr["pos"].T)))
            c.rdata.copy_(torch.from_numpy(np.ascontiguousarray(arr["rdata"].T)))
            c.ids.copy_(torch.from_numpy(arr["id"].astype(np.int64)))
        return c


class MultiFab:
    """The grid `state` (main.cpp:78): one periodic fab with ngrow ghost cells,
    [comp][k][j][i], rho/T/Ye set and ghosts filled (main.cpp:81-85)."""

    def __init__(self, emu: Emu):
        self.emu = emu
        p = emu.parms
        self.ng = emu.ngrow()
        n = p.ncell
        self.shape = (emu.ncomp, n[2] + 2 * self.ng[2], n[1] + 2 * self.ng[1], n[0] + 2 * self.ng[0])
        self.data = torch.zeros(self.shape, dtype=torch.float64, device=emu.device)
        self.data[0] = p.rho_g_ccm
        self.data[1] = p.T_MeV
        self.data[2] = p.Ye

    def view(self) -> capi.EmuGrid:
        g = capi.EmuGrid()
        g.data = self.data.data_ptr()
        n = self.emu.parms.ncell
        g.lo[:] = [-self.ng[0], -self.ng[1], -self.ng[2]]
        g.hi[:] = [n[0] - 1 + self.ng[0], n[1] - 1 + self.ng[1], n[2] - 1 + self.ng[2]]
        g.ncomp = self.emu.ncomp
        return g

    def valid(self) -> torch.Tensor:
        g, n = self.ng, self.emu.parms.ncell
        return self.data[:, g[2]:g[2] + n[2], g[1]:g[1] + n[1], g[0]:g[0] + n[0]]

    def FillBoundary(self):
        v = self.view()
        check(lib().emu_fill_boundary(self.emu.ctx, ctypes.byref(v)), "FillBoundary", self.emu.ctx)


def deposit_to_mesh(neutrinos: FlavoredNeutrinoContainer, state: MultiFab):
    """Evolve.cpp:91-126."""
    p, g = neutrinos.view(), state.view()
    check(lib().emu_deposit_to_mesh(state.emu.ctx, ctypes.byref(p), ctypes.byref(g)), "deposit_to_mesh", state.emu.ctx)


def interpolate_rhs_from_mesh(neutrinos_rhs: FlavoredNeutrinoContainer, state: MultiFab):
    """Evolve.cpp:128-176 (the rhs container holds a copy of the stage state)."""
    p, g = neutrinos_rhs.view(), state.view()
    check(lib().emu_interpolate_rhs_from_mesh(state.emu.ctx, ctypes.byref(p), ctypes.byref(g)),
          "interpolate_rhs_from_mesh", state.emu.ctx)


def compute_dt(state: MultiFab, Vvac_max: float) -> float:
    """Evolve.cpp:22-89 (cfl factors from the context's parameters)."""
    g = state.view()
    dt = ctypes.c_double()
    check(lib().emu_compute_dt(state.emu.ctx, ctypes.byref(g), Vvac_max, ctypes.byref(dt)), "compute_dt",
          state.emu.ctx)
    return dt.value


class Integrator:
    """The AMReX TimeIntegrator plus the post-step hook of main.cpp:155-199,
    on per-function calls (the plain drop-in path).  The tableau is the one
    the deck's integration.type / integration.rk.type select (classical RK4
    in every reference deck; forward Euler, trapezoid and SSPRK3 for the
    convergence study, Scripts/tests/convergence.sh)."""

    def __init__(self, emu: Emu, old: FlavoredNeutrinoContainer):
        self.emu = emu
        self.A, self.B = emu.parms.tableau()
        self.old = old
        self.new = FlavoredNeutrinoContainer(emu, old.np)
        self.F = [FlavoredNeutrinoContainer(emu, old.np) for _ in range(len(self.B))]
        self.state = MultiFab(emu)
        self.state.FillBoundary()
        self.vvac = old.Vvac_max()
        self.time = 0.0
        self.step_number = 0
        deposit_to_mesh(old, self.state)                  # main.cpp:118
        self.dt = compute_dt(self.state, self.vvac)        # main.cpp:206

    def rhs(self, F: FlavoredNeutrinoContainer, S: FlavoredNeutrinoContainer):
        deposit_to_mesh(S, self.state)                     # main.cpp:136
        self.state.FillBoundary()                          # main.cpp:137
        F.copyParticles(S)                                 # main.cpp:148
        interpolate_rhs_from_mesh(F, self.state)           # main.cpp:151

    def advance(self):
        # each stage state and the result: copyParticles(old), then one
        # ApplyFlavoredNeutrinoRHS per earlier stage -- one pass (rk_combine)
        dt = self.dt
        for i in range(len(self.B)):
            self.new.rk_combine(self.old, [dt * a for a in self.A[i]], self.F[:len(self.A[i])])
            self.rhs(self.F[i], self.new)
        self.new.rk_combine(self.old, [dt * b for b in self.B], self.F[:len(self.B)])
        self.time += dt
        self.step_number += 1
        # post_timestep_fun, main.cpp:155-199
        n = self.new
        n.SyncLocation(0)
        n.RedistributeLocal()
        n.SyncLocation(1)
        n.Renormalize()
        self.dt = compute_dt(self.state, self.vvac)
        self.old, self.new = self.new, self.old

    def integrate(self, nsteps: int, end_time: float = float("inf")):
        for _ in range(nsteps):
            if self.time >= end_time:
                break
            self.advance()
        return self.old


class FusedState:
    """emu_state: the fused, cell-sorted RK4 step (plain FP64 or FP32-bulk).
    slots: "auto", "packed" (cells' particles back to back) or "aligned" (each
    cell's run starts on a 128-byte boundary); results are the same."""

    def __init__(self, emu: Emu, np_: int, precision: str = "fp64", slots: str = "auto"):
        from .params import TABLEAUS
        if emu.parms.tableau() != TABLEAUS[4]:
            raise ValueError("the fused step is classical RK4 (integration.type = 1, integration.rk.type = 4); "
                             "use Integrator for the other tableaus")
        self.emu = emu
        self.np = int(np_)
        self.precision = {"fp64": capi.PLAIN_FP64, "fp32": capi.FP32_BULK}[precision]
        layout = {"auto": 0, "packed": capi.SLOTS_PACKED, "aligned": capi.SLOTS_ALIGNED}[slots]
        st = ctypes.c_void_p()
        check(lib().emu_state_create(emu.ctx, self.np, self.precision | layout, ctypes.byref(st)), "emu_state_create",
              emu.ctx)
        self.st = st

    def close(self):
        if getattr(self, "st", None):
            lib().emu_state_destroy(self.st)
            self.st = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, c: FlavoredNeutrinoContainer):
        v = c.view()
        check(lib().emu_state_load(self.st, ctypes.byref(v), ctypes.c_void_p(c.ids.data_ptr() if c.np else None)),
              "emu_state_load", self.emu.ctx)

    def set_vvac_max(self, vvac_max: float):
        """Replace the Vvac_max the load computed and recompute the starting dt
        (before the first evolve).  A run restarted from a checkpoint passes 0:
        the reference sets Vvac_max only when it initialises particles
        (FlavoredNeutrinoContainerInit.cpp:657-660)."""
        check(lib().emu_state_set_vvac_max(self.st, float(vvac_max)), "emu_state_set_vvac_max", self.emu.ctx)

    def store(self, c: FlavoredNeutrinoContainer):
        v = c.view()
        check(lib().emu_state_store(self.st, ctypes.byref(v), ctypes.c_void_p(c.ids.data_ptr() if c.np else None)),
              "emu_state_store", self.emu.ctx)

    def evolve(self, nsteps: int, fixed_dt: float = 0.0):
        check(lib().emu_state_evolve(self.st, int(nsteps), float(fixed_dt)), "emu_state_evolve", self.emu.ctx)

    def info(self, raise_on_error: bool = True):
        t, dt = ctypes.c_double(), ctypes.c_double()
        n, bad = ctypes.c_int64(), ctypes.c_int64()
        code = lib().emu_state_info(self.st, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n), ctypes.byref(bad))
        if raise_on_error:
            check(code, "emu_state (deferred)", self.emu.ctx, bad.value)
        return {"time": t.value, "dt": dt.value, "steps": n.value, "code": code, "bad_index": bad.value}

    def flavor_amplitude(self):
        """(sum over cells |N_01|, sum over cells |Nbar_01|)."""
        out = (ctypes.c_double * 2)()
        check(lib().emu_state_flavor_amplitude(self.st, out), "emu_state_flavor_amplitude", self.emu.ctx)
        return out[0], out[1]

    def profile(self, enable: bool = True):
        check(lib().emu_state_profile(self.st, int(enable)), "emu_state_profile", self.emu.ctx)

    def kernel_ms(self, k: int = 0):
        ms, n = ctypes.c_double(), ctypes.c_int64()
        check(lib().emu_state_kernel_ms(self.st, k, ctypes.byref(ms), ctypes.byref(n)), "emu_state_kernel_ms",
              self.emu.ctx)
        return ms.value, n.value


def evolve_flavor(parms: TestParams, nsteps: int | None = None, fused: bool = True, precision: str = "fp64"):
    """evolve_flavor (main.cpp:37-227) without plotfiles: initialise, then
    integrate nsteps RK4 steps; returns the final container."""
    emu = Emu(parms)
    parts = FlavoredNeutrinoContainer.init_particles(emu)
    n = parms.nsteps if nsteps is None else nsteps
    if fused:
        st = FusedState(emu, parts.np, precision)
        st.load(parts)
        st.evolve(n)
        st.store(parts)
        st.info()
        return parts
    return Integrator(emu, parts).integrate(n, parms.end_time)
