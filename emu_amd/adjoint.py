"""Reverse-mode gradients of the flavor evolution through a hand-written adjoint.

``FlavorObjective`` is a ``torch.autograd.Function``: its forward advances the
particles ``nsteps`` classical RK4 steps (the same stage algebra as the AMReX
RK4 integrator the reference uses, main.cpp:129-213) and returns the final
volume-integrated |N_emu|^2; its backward runs the discrete adjoint of those
steps.  Every right-hand side and its vector-Jacobian product run in
libemu_amd.so (include/emu_adjoint.h): the backward of the deposit is a
gather, the backward of the gather is a scatter, and the commutator has its
own adjoint.  Forward states are kept at checkpoints (every
``checkpoint_every`` steps; ``sqrt(nsteps)`` by default) and recomputed
segment by segment, either keeping the segment's step states
(``segment="store"``: (nsteps/cp + cp) states in memory) or recomputing each
step from the segment's start (``segment="recompute"``: nsteps/cp + 1 states,
about cp/2 more forward steps per step).  The footprint is checked against
``memory_budget`` bytes before anything runs.

Each forward step is one library call (emu_adj_step: one cell list, the four
stages with their deposits, Renormalize with a deferred status, the position
update) and each reverse step another (emu_adj_step_vjp: the stage states
recomputed, then the four stage VJPs).

Conventions: dt is held fixed over the run (compute_dt's maximum is not
differentiated); Renormalize is applied in the forward and passed straight
through in the backward (exact whenever its corrections stay below maxError,
as they do in the linear regime); positions do not depend on the flavour
state and carry no adjoint.
"""
from __future__ import annotations

import ctypes
import math

import torch

from .capi import check, lib


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


class AdjointProblem:
    """Fixed-order particle data and scratch for forward/adjoint sweeps."""

    def __init__(self, emu, parts, dt: float, nsteps: int, checkpoint_every: int | None = None,
                 segment: str = "store", memory_budget: float | None = None):
        self.emu = emu
        self.np = parts.np
        self.nh = emu.nh
        self.dt = float(dt)
        self.nsteps = int(nsteps)
        self.cp = int(checkpoint_every or max(1, int(math.sqrt(max(nsteps, 1)))))
        if segment not in ("store", "recompute"):
            raise ValueError("segment is 'store' or 'recompute'")
        self.segment = segment
        need = self.footprint_bytes()
        if memory_budget is not None and need > memory_budget:
            from .capi import EmuError
            raise EmuError(13, "AdjointProblem",
                           f"the checkpoints and work arrays need {need / 2**30:.1f} GiB, above the budget of "
                           f"{memory_budget / 2**30:.1f} GiB (fewer checkpoints or segment='recompute')")
        K = emu.K
        r = parts.rdata
        self.pos0 = parts.pos.clone()
        self.pup = r[K["pupx"]:K["pupt"] + 1].clone()
        self.Nn_base = torch.stack([r[K["N"]], r[K["Nbar"]]]).clone()
        self.Ll = torch.stack([r[K["L"]], r[K["Lbar"]]]).clone()
        f = K["f00_Re"]
        fb = K["f00_Rebar"]
        self.f0 = torch.cat([r[f:f + self.nh], r[fb:fb + self.nh]]).clone()
        nwork = lib().emu_adj_work_doubles(emu.ctx, self.np)
        self.work = torch.empty(nwork, dtype=torch.float64, device=emu.device)
        self.step_work = torch.empty(lib().emu_adj_step_work_doubles(emu.ctx, self.np), dtype=torch.float64,
                                     device=emu.device)
        self.lam_nbar = torch.zeros(self.np, dtype=torch.float64, device=emu.device)

    def footprint_bytes(self) -> float:
        """Device bytes of a gradient: the run's checkpoints (flavour state and
        positions), the segment's step states (segment='store'), the adjoint
        and its next value, the particle constants and the work arrays."""
        nh, np_ = self.nh, self.np
        state = 8.0 * np_ * (2 * nh + 3)
        ncp = -(-self.nsteps // self.cp)
        held = ncp + (self.cp if self.segment == "store" else 1)
        ncell = 1
        for n in self.emu.parms.ncell:
            ncell *= n
        work = 8.0 * (5 * ncell * 4 * nh + 11 * nh * np_ + 2 * np_) + 8.0 * (3 * ncell * 4 * nh + nh * np_ + np_)
        consts = 8.0 * np_ * (3 + 4 + 2 + 2 + 1 + 2 * nh)
        return held * state + 2 * 8.0 * 2 * nh * np_ + consts + work

    # ---- kernels ---------------------------------------------------------
    def rhs(self, pos, Nn, Y):
        k = torch.empty_like(Y)
        check(lib().emu_adj_rhs(self.emu.ctx, self.np, _p(pos), _p(self.pup), _p(Nn), _p(Y), _p(k), _p(self.work)),
              "emu_adj_rhs", self.emu.ctx)
        return k

    def vjp(self, pos, Nn, Y, lam_k):
        lam_Y = torch.empty_like(Y)
        check(lib().emu_adj_rhs_vjp(self.emu.ctx, self.np, _p(pos), _p(self.pup), _p(Nn), _p(Y), _p(lam_k),
                                    _p(lam_Y), _p(self.lam_nbar), _p(self.work)), "emu_adj_rhs_vjp", self.emu.ctx)
        return lam_Y

    def axpy(self, a, x, y, out=None):
        out = torch.empty_like(y) if out is None else out
        check(lib().emu_adj_axpy(self.emu.ctx, x.numel(), float(a), _p(x), _p(y), _p(out)), "emu_adj_axpy",
              self.emu.ctx)
        return out

    def objective(self, pos, Nn, Y, grad=True):
        J = ctypes.c_double()
        dJ = torch.empty_like(Y) if grad else None
        check(lib().emu_adj_objective(self.emu.ctx, self.np, _p(pos), _p(Nn), _p(Y), ctypes.byref(J),
                                      _p(dJ) if grad else None, _p(self.work)), "emu_adj_objective", self.emu.ctx)
        return J.value, dJ

    def advance_positions(self, pos):
        check(lib().emu_adj_advance_positions(self.emu.ctx, self.np, _p(pos), _p(self.pup), self.dt),
              "emu_adj_advance_positions", self.emu.ctx)

    def renormalize(self, Y):
        bad = ctypes.c_int64(-1)
        code = lib().emu_adj_renormalize(self.emu.ctx, self.np, _p(Y), _p(self.Ll), ctypes.byref(bad))
        check(code, "emu_adj_renormalize", self.emu.ctx, bad.value)

    # ---- one RK4 step (AMReX RK4 tableau, FlavoredNeutrinoContainer.H:39-46 saxpys)
    def stages(self, pos, Nn, y):
        dt = self.dt
        k1 = self.rhs(pos, Nn, y)
        Y2 = self.axpy(dt * 0.5, k1, y)
        k2 = self.rhs(pos, Nn, Y2)
        Y3 = self.axpy(dt * 0.5, k2, y)
        k3 = self.rhs(pos, Nn, Y3)
        Y4 = self.axpy(dt * 1.0, k3, y)
        return (y, Y2, Y3, Y4), (k1, k2, k3)

    def step_unfused(self, pos, Nn, y):
        """The same step from the per-RHS calls (kept as the check of step)."""
        dt = self.dt
        (Y1, Y2, Y3, Y4), (k1, k2, k3) = self.stages(pos, Nn, y)
        k4 = self.rhs(pos, Nn, Y4)
        acc = self.axpy(dt * (1. / 6.), k1, y)
        self.axpy(dt * (1. / 3.), k2, acc, acc)
        self.axpy(dt * (1. / 3.), k3, acc, acc)
        self.axpy(dt * (1. / 6.), k4, acc, acc)
        self.renormalize(acc)
        self.advance_positions(pos)
        return acc

    def step(self, pos, Nn, y):
        """y_{n+1} from y_n; pos advanced in place; Renormalize's status is
        checked by check_status()."""
        out = torch.empty_like(y)
        check(lib().emu_adj_step(self.emu.ctx, self.np, _p(pos), _p(self.pup), _p(Nn), _p(self.Ll), _p(y), _p(out),
                                 self.dt, _p(self.step_work)), "emu_adj_step", self.emu.ctx)
        return out

    def check_status(self):
        bad = ctypes.c_int64(-1)
        code = lib().emu_adj_status(self.emu.ctx, ctypes.byref(bad))
        check(code, "emu_adj_step (deferred Renormalize)", self.emu.ctx, bad.value)

    def step_adjoint(self, pos, Nn, y, lam):
        """lam_n from lam_{n+1} for the step that started at (pos, y):
        lambda_k4 = dt/6 lam, lambda_k3 = dt/3 lam + dt lY4, lambda_k2 =
        dt/3 lam + dt/2 lY3, lambda_k1 = dt/6 lam + dt/2 lY2, lam_n = lam +
        sum_i lY_i (in emu_adj_step_vjp)."""
        out = torch.empty_like(lam)
        check(lib().emu_adj_step_vjp(self.emu.ctx, self.np, _p(pos), _p(self.pup), _p(Nn), _p(y), _p(lam), _p(out),
                                     _p(self.lam_nbar), self.dt, _p(self.step_work)), "emu_adj_step_vjp",
              self.emu.ctx)
        return out

    # ---- whole runs ------------------------------------------------------
    def Nn(self, ratio: float):
        Nn = self.Nn_base.clone()
        Nn[1] *= ratio
        return Nn

    def run(self, f0, ratio: float, keep_checkpoints=True):
        Nn = self.Nn(ratio)
        pos = self.pos0.clone()
        y = f0.clone()
        cps = []
        for n in range(self.nsteps):
            if keep_checkpoints and n % self.cp == 0:
                cps.append((n, pos.clone(), y.clone()))
            y = self.step(pos, Nn, y)
        J, dJ = self.objective(pos, Nn, y, grad=keep_checkpoints)
        self.check_status()
        return J, dJ, cps, Nn

    def adjoint_sweep(self, lam, cps, Nn):
        """Backward through the recorded run: lam = dJ/dy_N -> dJ/dy_0; the
        nubar-weight adjoint accumulates in self.lam_nbar."""
        self.lam_nbar.zero_()
        for (n0, pos0, y0) in reversed(cps):
            n1 = min(n0 + self.cp, self.nsteps)
            if self.segment == "store":
                states = []  # the segment's step-start states, recomputed
                pos = pos0.clone()
                y = y0.clone()
                for n in range(n0, n1):
                    states.append((pos.clone(), y.clone()))
                    if n + 1 < n1:
                        y = self.step(pos, Nn, y)
                for (pos_n, y_n) in reversed(states):
                    lam = self.step_adjoint(pos_n, Nn, y_n, lam)
            else:
                for n in range(n1 - 1, n0 - 1, -1):  # each step's start, from the checkpoint
                    pos = pos0.clone()
                    y = y0
                    for _ in range(n0, n):
                        y = self.step(pos, Nn, y)
                    lam = self.step_adjoint(pos, Nn, y, lam)
        self.check_status()
        return lam

    def gradients(self, f0, ratio: float):
        """J, dJ/df0 and dJ/dratio by the adjoint sweep."""
        J, lam, cps, Nn = self.run(f0, ratio)
        lam = self.adjoint_sweep(lam, cps, Nn)
        dratio = float((self.lam_nbar * self.Nn_base[1]).sum().item())
        return J, lam, dratio


class FlavorObjective(torch.autograd.Function):
    """J(f0, ratio) = final sum over cells |N_emu|^2 after nsteps RK4 steps,
    with nubar weights scaled by `ratio`."""

    @staticmethod
    def forward(ctx, f0, ratio, problem: AdjointProblem):
        J, lam, cps, Nn = problem.run(f0.detach(), float(ratio))
        ctx.problem = problem
        ctx.saved = (lam, cps, Nn)
        return torch.tensor(J, dtype=torch.float64, device=f0.device)

    @staticmethod
    def backward(ctx, gJ):
        p = ctx.problem
        lam, cps, Nn = ctx.saved
        grad_f0 = p.adjoint_sweep(lam, cps, Nn)
        grad_ratio = (p.lam_nbar * p.Nn_base[1]).sum()
        ctx.saved = None
        return gJ * grad_f0, gJ * grad_ratio, None
