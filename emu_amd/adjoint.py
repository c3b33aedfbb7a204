"""Reverse-mode gradients of the flavor evolution through a hand-written adjoint.

``FlavorObjective`` is a ``torch.autograd.Function``: its forward advances the
particles ``nsteps`` classical RK4 steps (the same stage algebra as the AMReX
RK4 integrator the reference uses, main.cpp:129-213) and returns the final
volume-integrated |N_emu|^2; its backward runs the discrete adjoint of those
steps.  Every right-hand side and its vector-Jacobian product run in
libemu_amd.so (include/emu_adjoint.h): the backward of the deposit is a
gather, the backward of the gather is a scatter, and the commutator has its
own adjoint.  Forward states are kept at checkpoints and recomputed segment by
segment (every ``checkpoint_every`` steps; ``sqrt(nsteps)`` by default).

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

    def __init__(self, emu, parts, dt: float, nsteps: int, checkpoint_every: int | None = None):
        self.emu = emu
        self.np = parts.np
        self.nh = emu.nh
        self.dt = float(dt)
        self.nsteps = int(nsteps)
        self.cp = int(checkpoint_every or max(1, int(math.sqrt(max(nsteps, 1)))))
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
        self.lam_nbar = torch.zeros(self.np, dtype=torch.float64, device=emu.device)

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

    def step(self, pos, Nn, y):
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

    def step_adjoint(self, pos, Nn, y, lam):
        """lam_n from lam_{n+1} for the step that started at (pos, y)."""
        dt = self.dt
        (Y1, Y2, Y3, Y4), _ = self.stages(pos, Nn, y)
        zero = torch.zeros_like(lam)
        # lambda_k4 = dt/6 lam; lambda_k3 = dt/3 lam + dt lY4; lambda_k2 = dt/3 lam + dt/2 lY3;
        # lambda_k1 = dt/6 lam + dt/2 lY2; lam_n = lam + sum_i lY_i
        lY4 = self.vjp(pos, Nn, Y4, self.axpy(dt * (1. / 6.), lam, zero))
        lY3 = self.vjp(pos, Nn, Y3, self.axpy(dt * 1.0, lY4, self.axpy(dt * (1. / 3.), lam, zero)))
        lY2 = self.vjp(pos, Nn, Y2, self.axpy(dt * 0.5, lY3, self.axpy(dt * (1. / 3.), lam, zero)))
        lY1 = self.vjp(pos, Nn, Y1, self.axpy(dt * 0.5, lY2, self.axpy(dt * (1. / 6.), lam, zero)))
        out = self.axpy(1.0, lY1, lam)
        for l in (lY2, lY3, lY4):
            self.axpy(1.0, l, out, out)
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
        return J, dJ, cps, Nn

    def adjoint_sweep(self, lam, cps, Nn):
        """Backward through the recorded run: lam = dJ/dy_N -> dJ/dy_0; the
        nubar-weight adjoint accumulates in self.lam_nbar."""
        self.lam_nbar.zero_()
        for (n0, pos0, y0) in reversed(cps):
            n1 = min(n0 + self.cp, self.nsteps)
            states = []  # the segment's step-start states, recomputed
            pos = pos0.clone()
            y = y0.clone()
            for n in range(n0, n1):
                states.append((pos.clone(), y.clone()))
                if n + 1 < n1:
                    y = self.step(pos, Nn, y)
            for (pos_n, y_n) in reversed(states):
                lam = self.step_adjoint(pos_n, Nn, y_n, lam)
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
