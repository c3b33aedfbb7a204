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
segment by segment, either keeping the segment's step states with each
step's stage states and deposits (``segment="keep"``, the default: the
reverse step then recomputes nothing), keeping the step states only
(``segment="store"``: (nsteps/cp + cp) states in memory; each reverse step
recomputes its stages) or recomputing each step from the segment's start
(``segment="recompute"``: nsteps/cp + 1 states, about cp/2 more forward
steps per step), or reversing each segment on a binomial schedule with
``snapshots`` more step states (``segment="binomial"``, Griewank's revolve:
t steps with s snapshots cost about r t forward steps, r the least count
with C(s + r, s) >= t, instead of t^2/2).  The footprint is checked against
``memory_budget`` bytes before anything runs.

Each forward step is one library call (emu_adj_step: one cell list, the four
stages with their deposits, Renormalize with a deferred status, the position
update) and each reverse step another (emu_adj_step_vjp: the stage states
recomputed, or taken from the kept ones, then the four stage VJPs).

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


def _take(x):
    """x, or the one tensor a one-element list holds (the list is emptied).
    A call's arguments stay referenced by the caller until the call returns,
    so a large array handed over in a list is the only way the callee can
    free it early."""
    return x.pop() if isinstance(x, list) else x


def binomial_split(t: int, s: int) -> int:
    """Where the binomial schedule stores its next snapshot when reversing t
    steps from a held state with s free snapshot slots: after m steps, so that
    the t - m steps to its right are reversed from the new snapshot with s - 1
    free slots and the m to its left from the held state with s.  With the
    held state counted, S = s + 1 states reverse at most b(S, r) = C(S + r, S)
    steps when no step runs forward more than r times (Griewank, Algorithm
    799: revolve, ACM TOMS 26, 2000), at the least cost r t - C(S + r, r - 1)
    for the least such r.  The split reaches that cost when the right part
    fits S - 1 states in r runs but needs more than r - 1 (t - m <= b(S - 1, r),
    t - m >= b(S - 1, r - 1)) and the left part fits S states in r - 1 runs
    but needs more than r - 2 (m <= b(S, r - 1), m >= b(S, r - 2)); the
    smallest such m is taken."""
    if t < 2 or s < 1:
        raise ValueError("binomial_split needs t >= 2 steps and s >= 1 slots")
    S = s + 1

    def b(k, r):
        return math.comb(k + r, k) if r >= 0 else 0

    r = 1
    while b(S, r) < t:
        r += 1
    return max(1, t - b(S - 1, r), b(S, r - 2))


def binomial_reverse(start, n0: int, n1: int, s: int, advance, reverse):
    """Reverse the steps n0 .. n1-1, last first, from `start`, the state at
    step n0, holding at most s more states at once.  advance(state, n) returns
    the state at step n >= state's step, run forward from `state` (a new
    state: `state` stays valid); reverse(state_at_n, n) goes back through
    step n.  With s = 0 every step is run again from `start`."""
    while n1 > n0:
        t = n1 - n0
        if t == 1 or s == 0:
            for n in range(n1 - 1, n0 - 1, -1):
                reverse(start if n == n0 else advance(start, n), n)
            return
        m = n0 + binomial_split(t, s)
        snap = advance(start, m)
        binomial_reverse(snap, m, n1, s - 1, advance, reverse)
        del snap
        n1 = m


class AdjointProblem:
    """Fixed-order particle data and scratch for forward/adjoint sweeps."""

    def __init__(self, emu, parts, dt: float, nsteps: int, checkpoint_every: int | None = None,
                 segment: str = "keep", memory_budget: float | None = None, resort_every: int = 4,
                 precision: str = "fp64", snapshots: int = 2):
        self.emu = emu
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision is 'fp64' or 'fp32'")
        # fp32: the FP32-bulk form (float flavour arrays, particle constants and
        # adjoints; FP64 positions, deposit sums, Renormalize and lam_nbar)
        self.precision = precision
        self.dtype = torch.float64 if precision == "fp64" else torch.float32
        self._sfx = "" if precision == "fp64" else "_f32"
        self.np = parts.np
        self.nh = emu.nh
        self.dt = float(dt)
        self.nsteps = int(nsteps)
        self.cp = int(checkpoint_every or max(1, int(math.sqrt(max(nsteps, 1)))))
        # the particle arrays go back to cell order every `resort` steps, and at
        # every checkpoint (so resort divides cp)
        self.resort = math.gcd(self.cp, max(1, int(resort_every)))
        if segment not in ("keep", "store", "recompute", "binomial"):
            raise ValueError("segment is 'keep', 'store', 'recompute' or 'binomial'")
        self.segment = segment
        # binomial: step states held at once within a segment, beside its checkpoint
        self.snapshots = max(0, int(snapshots))
        self.keep_doubles = int(self._fn("emu_adj_step_keep_doubles")(emu.ctx, self.np))
        self.step_work_doubles = int(self._fn("emu_adj_step_work_doubles")(emu.ctx, self.np))
        self.work_doubles = int(lib().emu_adj_work_doubles(emu.ctx, self.np)) if precision == "fp64" else 0
        self._keeps = None
        self.steps_run = 0  # forward steps run so far: the runs' own and the sweeps' replays
        need = self.footprint_bytes()
        if memory_budget is not None and need > memory_budget:
            from .capi import EmuError
            raise EmuError(13, "AdjointProblem",
                           f"the checkpoints and work arrays need {need / 2**30:.1f} GiB, above the budget of "
                           f"{memory_budget / 2**30:.1f} GiB (fewer checkpoints, or segment='recompute' or "
                           f"'binomial' with fewer snapshots)")
        K = emu.K
        r = parts.rdata
        dt_ = self.dtype
        self.pos0 = parts.pos.clone()
        self.pup = r[K["pupx"]:K["pupt"] + 1].to(dt_, copy=True)
        Nn = torch.stack([r[K["N"]], r[K["Nbar"]]])
        # fp32: the weights times a power of two that puts the largest in [1, 2)
        # (FP32 cannot hold N = n dV of large cells, nor the objective's
        # adjoints, which grow as N^2); the library takes it back out of the
        # evolution, and J and its gradients come back divided by its square
        self.wscale = 1.0
        if precision == "fp32" and self.np > 0:
            nmax = float(Nn.abs().max())
            if nmax > 0 and math.isfinite(nmax):
                self.wscale = 2.0 ** -math.floor(math.log2(nmax))
        self.Nn_base = (Nn * self.wscale).to(dt_)
        self.Ll = torch.stack([r[K["L"]], r[K["Lbar"]]]).to(dt_)
        f = K["f00_Re"]
        fb = K["f00_Rebar"]
        self.f0 = torch.cat([r[f:f + self.nh], r[fb:fb + self.nh]]).to(dt_)
        # the per-RHS calls' work (FP64 only); the objective and the step calls use step_work
        self.work = (torch.empty(self.work_doubles, dtype=torch.float64, device=emu.device)
                     if precision == "fp64" else None)
        self.step_work = torch.empty(self.step_work_doubles, dtype=torch.float64, device=emu.device)
        self.lam_nbar = torch.zeros(self.np, dtype=torch.float64, device=emu.device)

    def _fn(self, name):
        return getattr(lib(), name + self._sfx)

    def footprint_bytes(self) -> float:
        """Device bytes at the peak of a gradient.  Held throughout: the work
        arrays, the kept stage data, the particle constants, the run's
        checkpoints and the segment's held step states (a state: flavour
        state, FP64 positions, cell order; the binomial schedule's snapshots,
        or a stored segment's steps).  On top, the larger of the two phases:
          - a reverse step: the state it goes back through, the adjoint and
            its next value, the step's constants in its order, and the index
            arrays of the adjoint's move between orders;
          - a forward step across a re-sort (in the run or a replay): the
            adjoint, the flavour state and its next or re-sorted copy, two
            copies of the positions, the old and new constants.
        On the 128^3 box (FP32-bulk, 100 steps, two binomial snapshots) this
        says 184 GiB against 178 GiB measured at the peak."""
        nh2, np_ = 2 * self.nh, self.np
        e = 8.0 if self.precision == "fp64" else 4.0
        flav = e * nh2 * np_
        pos = 8.0 * 3 * np_
        state = flav + pos + 4.0 * np_
        ncp = -(-self.nsteps // self.cp)
        seg = {"recompute": 0, "binomial": self.snapshots}.get(self.segment, self.cp)
        kept = 8.0 * self.cp * self.keep_doubles if self.segment == "keep" else 0.0
        work = 8.0 * (self.step_work_doubles + self.work_doubles)
        step_consts = e * 8 * np_  # pup, Nn, Ll in a step's order
        # a stored or kept segment holds its steps' constants, one set per re-sort
        if self.segment in ("keep", "store"):
            kept += -(-self.cp // self.resort) * step_consts
        # pos0; pup, Nn, Ll and Nn(ratio) in the problem's order; f0 while held; lam_nbar twice
        f0 = flav if getattr(self, "f0", True) is not None else 0.0
        base = work + kept + (ncp + seg) * state + pos + e * 10 * np_ + f0 + 16.0 * np_
        reverse = state + 2 * flav + step_consts + 24.0 * np_
        forward = 3 * flav + 2 * pos + 2 * step_consts
        return base + max(reverse, forward)

    # ---- kernels ---------------------------------------------------------
    def _fp64_only(self, what):
        if self.precision != "fp64":
            raise NotImplementedError(what + " runs in FP64 only (precision='fp64')")

    def rhs(self, pos, Nn, Y):
        self._fp64_only("rhs")
        k = torch.empty_like(Y)
        check(lib().emu_adj_rhs(self.emu.ctx, self.np, _p(pos), _p(self.pup), _p(Nn), _p(Y), _p(k), _p(self.work)),
              "emu_adj_rhs", self.emu.ctx)
        return k

    def vjp(self, pos, Nn, Y, lam_k):
        self._fp64_only("vjp")
        lam_Y = torch.empty_like(Y)
        check(lib().emu_adj_rhs_vjp(self.emu.ctx, self.np, _p(pos), _p(self.pup), _p(Nn), _p(Y), _p(lam_k),
                                    _p(lam_Y), _p(self.lam_nbar), _p(self.work)), "emu_adj_rhs_vjp", self.emu.ctx)
        return lam_Y

    def axpy(self, a, x, y, out=None):
        self._fp64_only("axpy")
        out = torch.empty_like(y) if out is None else out
        check(lib().emu_adj_axpy(self.emu.ctx, x.numel(), float(a), _p(x), _p(y), _p(out)), "emu_adj_axpy",
              self.emu.ctx)
        return out

    def objective(self, pos, Nn, Y, grad=True):
        J = ctypes.c_double()
        dJ = torch.empty_like(Y) if grad else None
        check(self._fn("emu_adj_objective")(self.emu.ctx, self.np, _p(pos), _p(Nn), _p(Y), ctypes.byref(J),
                                            _p(dJ) if grad else None, _p(self.step_work)), "emu_adj_objective",
              self.emu.ctx)
        return J.value, dJ

    def advance_positions(self, pos):
        self._fp64_only("advance_positions")
        check(lib().emu_adj_advance_positions(self.emu.ctx, self.np, _p(pos), _p(self.pup), self.dt),
              "emu_adj_advance_positions", self.emu.ctx)

    def renormalize(self, Y):
        self._fp64_only("renormalize")
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

    def step(self, pos, Nn, y, consts=None, keep=None):
        """y_{n+1} from y_n; pos advanced in place; Renormalize's status is
        checked by check_status().  consts = (pup, Ll) in the arrays' order
        (the problem's own order when None).  With keep (keep_doubles
        doubles), the step's stage states and deposits stay there for
        step_adjoint."""
        pup, Ll = consts if consts is not None else (self.pup, self.Ll)
        out = torch.empty_like(y)
        ws = (self.wscale,) if self.precision == "fp32" else ()
        self.steps_run += 1
        if keep is None:
            check(self._fn("emu_adj_step")(self.emu.ctx, self.np, _p(pos), _p(pup), _p(Nn), _p(Ll), _p(y), _p(out),
                                           self.dt, *ws, _p(self.step_work)), "emu_adj_step", self.emu.ctx)
        else:
            check(self._fn("emu_adj_step_keep")(self.emu.ctx, self.np, _p(pos), _p(pup), _p(Nn), _p(Ll), _p(y),
                                                _p(out), self.dt, *ws, _p(self.step_work), _p(keep)),
                  "emu_adj_step_keep", self.emu.ctx)
        return out

    def check_status(self):
        bad = ctypes.c_int64(-1)
        code = lib().emu_adj_status(self.emu.ctx, ctypes.byref(bad))
        check(code, "emu_adj_step (deferred Renormalize)", self.emu.ctx, bad.value)

    def step_adjoint(self, pos, Nn, y, lam, consts=None, keep=None):
        """lam_n from lam_{n+1} for the step that started at (pos, y):
        lambda_k4 = dt/6 lam, lambda_k3 = dt/3 lam + dt lY4, lambda_k2 =
        dt/3 lam + dt/2 lY3, lambda_k1 = dt/6 lam + dt/2 lY2, lam_n = lam +
        sum_i lY_i (in emu_adj_step_vjp).  lam_nbar accumulates in the
        arrays' order.  keep: the step's kept stage data (step(..., keep)),
        which spares the recompute."""
        pup = consts[0] if consts is not None else self.pup
        out = torch.empty_like(lam)
        ws = (self.wscale,) if self.precision == "fp32" else ()
        if keep is None:
            check(self._fn("emu_adj_step_vjp")(self.emu.ctx, self.np, _p(pos), _p(pup), _p(Nn), _p(y), _p(lam),
                                               _p(out), _p(self.lam_nbar), self.dt, *ws, _p(self.step_work)),
                  "emu_adj_step_vjp", self.emu.ctx)
        else:
            check(self._fn("emu_adj_step_vjp_kept")(self.emu.ctx, self.np, _p(pos), _p(pup), _p(Nn), _p(y), _p(keep),
                                                    _p(lam), _p(out), _p(self.lam_nbar), self.dt, *ws,
                                                    _p(self.step_work)), "emu_adj_step_vjp_kept", self.emu.ctx)
        return out

    # ---- cell order ------------------------------------------------------
    # The step calls read every particle's arrays through the cell list; as
    # the particles drift from the order they started in, those reads
    # scatter (a 200-step gradient ran at 0.45x the rate of a 20-step one).
    # So every `resort` steps the particle arrays go back to cell order.
    # `order` maps an array order to the problem's: element i of an array in
    # that order is particle order[i].  An adjoint carries the order it is in
    # and is moved into the order of each step it goes back through.
    def _permute(self, x, idx, mode, out=None):
        rows = x.numel() // self.np if self.np else 0
        if out is None:
            out = torch.empty_like(x) if mode == 0 else torch.zeros_like(x)
        fn = lib().emu_adj_permute_f32 if x.dtype == torch.float32 else lib().emu_adj_permute
        check(fn(self.emu.ctx, rows, self.np, _p(x), _p(idx), _p(out), mode), "emu_adj_permute", self.emu.ctx)
        return out

    def _cell_order(self, pos):
        o = torch.empty(self.np, dtype=torch.int32, device=pos.device)
        check(lib().emu_adj_cell_order(self.emu.ctx, self.np, _p(pos), _p(o), _p(self.step_work)),
              "emu_adj_cell_order", self.emu.ctx)
        return o

    def _segment_consts(self, order, Nn):
        """(pup, Nn, Ll) in the given order."""
        return (self._permute(self.pup, order, 0), self._permute(Nn, order, 0), self._permute(self.Ll, order, 0))

    def _resort(self, pos, y, order, Nn):
        """The arrays in cell order: (pos, y, order, consts)."""
        o = self._cell_order(pos)
        pos, y = self._permute(pos, o, 0), self._permute(y, o, 0)
        order = order[o.long()]
        return pos, y, order, self._segment_consts(order, Nn)

    def _replay(self, cp, Nn, upto):
        """The step-start states n0 .. upto of a checkpoint's segment, as the
        forward run made them (re-sorting at the same steps): yields
        (n, pos, y, order, consts); pos is advanced in place by the next
        step, so a caller that keeps it clones it."""
        n0, pos, y, order = cp
        pos = pos.clone()
        consts = self._segment_consts(order, Nn)
        for n in range(n0, upto + 1):
            if n > n0 and n % self.resort == 0:
                pos, y, order, consts = self._resort(pos, y, order, Nn)
            yield n, pos, y, order, consts
            if n < upto:
                y = self.step(pos, consts[1], y, (consts[0], consts[2]))

    def _advance(self, state, Nn, n):
        """The state (n, pos, y, order) at step n >= state's, run forward from
        `state` (which stays valid) with the forward run's re-sorts.  Unlike
        _replay it holds only the current step's arrays, so that a re-sort
        (a new copy of the positions and the flavour state) comes on top of
        one state, not two."""
        m, pos, y, order = state
        pos = pos.clone()
        consts = self._segment_consts(order, Nn)
        for m in range(m, n):
            y = self.step(pos, consts[1], y, (consts[0], consts[2]))
            if (m + 1) % self.resort == 0:
                consts = None
                pos, y, order, consts = self._resort(pos, y, order, Nn)
        return n, pos, y, order

    def _replay_keep(self, cp, Nn, n1):
        """The steps n0 .. n1-1 of a checkpoint's segment run again (re-sorting
        at the same steps as the forward run), each keeping its stage data:
        a list of (n, pos, y, order, consts, keep) at the steps' starts."""
        n0, pos, y, order = cp
        pos = pos.clone()
        consts = self._segment_consts(order, Nn)
        if self._keeps is None:
            self._keeps = [torch.empty(self.keep_doubles, dtype=torch.float64, device=pos.device)
                           for _ in range(self.cp)]
        out = []
        for i, n in enumerate(range(n0, n1)):
            if n > n0 and n % self.resort == 0:
                pos, y, order, consts = self._resort(pos, y, order, Nn)
            start = pos.clone()
            y_next = self.step(pos, consts[1], y, (consts[0], consts[2]), keep=self._keeps[i])
            out.append((n, start, y, order, consts, self._keeps[i]))
            y = y_next
        return out

    def _to_order(self, lam, lam_order, order):
        """lam from lam_order into order, in one gather: element i of the
        result is particle order[i], which lam holds where lam_order has it."""
        if order is lam_order:
            return lam
        inv = torch.empty_like(lam_order)
        inv[lam_order.long()] = torch.arange(self.np, dtype=lam_order.dtype, device=lam_order.device)
        return self._permute(lam, inv[order.long()], 0)

    # ---- whole runs ------------------------------------------------------
    def Nn(self, ratio: float):
        Nn = self.Nn_base.clone()
        Nn[1] *= ratio
        return Nn

    def run(self, f0, ratio: float, keep_checkpoints=True, consume=False):
        """J after nsteps from f0 (problem order); with keep_checkpoints, also
        dJ/dy_N and the checkpoints (n, pos, y, order) for the sweep, and
        the order dJ/dy_N is in.  consume: f0's memory becomes the run's state
        and the initial positions the run's positions (no copies; f0 then
        holds no meaningful values, and the problem runs no more); f0 given
        as a one-element list is taken out of it, so that it is freed once the
        run has moved on."""
        f0 = _take(f0)
        if self.pos0 is None:
            raise RuntimeError("AdjointProblem.run after a run with consume=True")
        Nn = self.Nn(ratio)
        pos = self.pos0.clone() if not consume else self.pos0
        if consume:
            self.pos0 = None
        y = f0 if consume else f0.clone()
        del f0
        order = torch.arange(self.np, dtype=torch.int32, device=y.device)
        cps = []
        for n in range(self.nsteps):
            if n % self.resort == 0:
                pos, y, order, consts = self._resort(pos, y, order, Nn)
            if keep_checkpoints and n % self.cp == 0:
                cps.append((n, pos.clone(), y.clone(), order))
            y = self.step(pos, consts[1], y, (consts[0], consts[2]))
        J, dJ = self.objective(pos, consts[1], y, grad=keep_checkpoints)
        self.check_status()
        return J / self.wscale**2, dJ, cps, Nn, order

    def adjoint_sweep(self, lam_end, cps, Nn, order_end, progress=None):
        """Backward through the recorded run: lam_end = dJ/dy_N (in order_end)
        -> dJ/dy_0 in the problem's order; the nubar-weight adjoint ends in
        self.lam_nbar, in the problem's order too.  lam_end given as a
        one-element list is taken out of it: then nothing else holds it, and
        it is freed after the first reverse step.  progress(n), when given, is
        called after the reverse step through step n."""
        lam_nbar = torch.zeros_like(self.lam_nbar)
        self.lam_nbar.zero_()
        # the adjoint lives in `lam` alone, which each reverse step rebinds (an
        # argument that a nested function rebinds stays referenced by the
        # frame, so the arguments are moved out first)
        lam, lam_order = _take(lam_end), order_end
        del lam_end, order_end
        nbar_order = lam_order

        def back_through(n, pos, y, order, consts, keep=None):
            # the sweep's adjoint into this step's order, then back through the
            # step; each rebinding drops the old array as soon as the new exists
            nonlocal lam, lam_order, nbar_order
            lam = self._to_order(lam, lam_order, order)
            lam_order = order
            if order is not nbar_order:
                self._permute(self.lam_nbar, nbar_order, 2, out=lam_nbar)
                self.lam_nbar.zero_()
                nbar_order = order
            lam = self.step_adjoint(pos, consts[1], y, lam, (consts[0], consts[2]), keep=keep)
            if progress is not None:
                progress(n)

        for cp in reversed(cps):
            n0 = cp[0]
            n1 = min(n0 + self.cp, self.nsteps)
            # (each branch lets go of its segment's states before the next segment's are made)
            if self.segment == "keep":
                for st in reversed(self._replay_keep(cp, Nn, n1)):
                    back_through(*st)
                del st
            elif self.segment == "store":
                states = [(n, pos.clone(), y, order, consts) for (n, pos, y, order, consts)
                          in self._replay(cp, Nn, n1 - 1)]
                for st in reversed(states):
                    back_through(*st)
                del states, st
            elif self.segment == "binomial":
                def reverse(state, n):
                    back_through(*state, self._segment_consts(state[3], Nn))

                binomial_reverse(cp, n0, n1, self.snapshots, lambda st, n: self._advance(st, Nn, n), reverse)
            else:
                for n in range(n1 - 1, n0 - 1, -1):  # each step's start, run again from the checkpoint
                    st = self._advance(cp, Nn, n)
                    back_through(*st, self._segment_consts(st[3], Nn))
                    del st
        self._permute(self.lam_nbar, nbar_order, 2, out=lam_nbar)
        self.check_status()
        self.lam_nbar.copy_(lam_nbar)
        lam = self._permute(lam, lam_order, 1)
        if self.wscale == 1.0:
            return lam
        # the unscaled gradient outgrows FP32's range: it comes back in FP64
        # (the power-of-two factor divides out exactly)
        lam = lam.double()
        return lam.mul_(1.0 / self.wscale**2)

    def ratio_gradient(self):
        """dJ/dratio from the nubar-weight adjoint of the last sweep."""
        return (self.lam_nbar * self.Nn_base[1].double()).sum() / self.wscale**2

    def gradients(self, f0, ratio: float):
        """J, dJ/df0 and dJ/dratio by the adjoint sweep."""
        J, lam, cps, Nn, order = self.run(f0, ratio)
        lam = [lam]  # the sweep holds dJ/dy_N alone
        lam = self.adjoint_sweep(lam, cps, Nn, order)
        return J, lam, float(self.ratio_gradient().item())


class FlavorObjective(torch.autograd.Function):
    """J(f0, ratio) = final sum over cells |N_emu|^2 after nsteps RK4 steps,
    with nubar weights scaled by `ratio`."""

    @staticmethod
    def forward(ctx, f0, ratio, problem: AdjointProblem):
        J, lam, cps, Nn, order = problem.run(f0.detach().to(problem.dtype), float(ratio))
        ctx.f0_dtype = f0.dtype
        ctx.problem = problem
        ctx.saved = (lam, cps, Nn, order)
        return torch.tensor(J, dtype=torch.float64, device=f0.device)

    @staticmethod
    def backward(ctx, gJ):
        p = ctx.problem
        saved = list(ctx.saved)
        ctx.saved = None
        cps, Nn, order = saved[1:]
        # the sweep holds dJ/dy_N alone, so that it goes as soon as the first step is through
        grad_f0 = p.adjoint_sweep([saved.pop(0)], cps, Nn, order)
        grad_ratio = p.ratio_gradient()
        return (gJ * grad_f0).to(ctx.f0_dtype), gJ * grad_ratio, None
