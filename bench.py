#!/usr/bin/env python3
"""Particle-updates/s of the fused RK4 flavor-evolution step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ff64]

One particle-update is one full RK4 step of one particle (four right-hand-side
evaluations with their deposits, the stage updates, Renormalize and the
periodic wrap; each particle carries nu and nubar) -- the reference's figure of
merit (Source/main.cpp:181,219) times 1e6.  The workload is the 3-flavour
fast-flavor box of 64^3 cells x 92 directions (24,117,248 particles).  Both
the plain FP64 step and the FP32-bulk step run K timed steps from the same
initial particles; `value` is the FP32-bulk rate, the FP64 rate and the
growth-rate error of |N_emu| between the two are reported beside it.

The timed region is the step loop only (inputs resident in HBM), bracketed by
stream synchronisation and HIP events on the library's stream; an event pair
around every launch gives the stage kernel's average for the roofline
(`roofline.frac`), and the whole step's wall time gives `roofline.wall_frac`.
After the timed runs, an untimed precision check runs FP32-bulk and FP64 for
2000 steps through saturation on the 1-D fiducial workload (growth rate of
|N_emu|, saturation, cell-averaged N_ab, trace drift).
The CPU baseline is the FP64 oracle (OpenMP, reference-structured AoS with the
same Copy/Saxpy RK4 pattern) on a 16^3-cell box of the same physics, timed on
this host's cores.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD public figure


def algorithmic_bytes(nf: int, ppc: float, precision: str) -> float:
    """Compulsory HBM bytes per particle-update of a 3-register classical RK4
    whose mesh coupling deposits before every gather: 65 + 20 S values per
    update (S = 2 NF^2 flavour reals), 27 of them positions kept in FP64, plus
    the grid term 12 G 8 / ppc (G = 8 NF^2 deposited components, FP64
    read-modify-write deposit and gather read over 4 stages)."""
    S = 2 * nf * nf
    G = 8 * nf * nf
    values = 65 + 20 * S
    part = 8 * values if precision == "fp64" else 8 * 27 + 4 * (values - 27)
    return part + 12 * G * 8 / ppc


def counter_traffic(workload: str, precision: str):
    """HBM bytes per k_stage launch measured with rocprofv3 --pmc (FETCH_SIZE
    and WRITE_SIZE passes, scaled by a 1 GiB copy) and committed as a text
    summary under profiles/; None if there is none for this workload."""
    path = os.path.join(ROOT, "profiles", f"k_stage_{workload}_traffic.txt")
    if not os.path.exists(path):
        return None
    key = f"{precision} traffic_bytes_per_launch ="
    for line in open(path):
        if line.startswith(key):
            return float(line.split("=", 1)[1].split()[0])
    return None


def run_precision(core, parms, parts, precision, steps, warmup):
    import torch
    emu = parts.emu
    st = core.FusedState(emu, parts.np, precision)
    st.load(parts)
    if warmup:
        st.evolve(warmup)
    info0 = st.info()
    a0 = st.flavor_amplitude()
    # the library launches on the context's stream, which core.Emu binds to
    # torch's current stream; events on that stream bracket the timed steps
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    # HIP events around every kernel launch of the timed steps, on the same stream
    st.profile(True)
    torch.cuda.synchronize()
    t_wall0 = time.perf_counter()
    ev0.record(stream)
    st.evolve(steps)
    ev1.record(stream)
    torch.cuda.synchronize()
    t_wall = time.perf_counter() - t_wall0
    ms = ev0.elapsed_time(ev1)
    stage_ms, n_stage = st.kernel_ms(0)
    fill_ms, _ = st.kernel_ms(1)
    sort_ms, _ = st.kernel_ms(2)
    dt_ms, _ = st.kernel_ms(3)
    st.profile(False)
    info1 = st.info()
    a1 = st.flavor_amplitude()
    rate = math.log(a1[0] / a0[0]) / (info1["time"] - info0["time"]) if a0[0] > 0 and a1[0] > 0 else float("nan")
    st.close()
    return {"ms_total": ms, "wall_s": t_wall, "ms_per_step": ms / steps, "updates_per_s": parts.np * steps / (ms * 1e-3),
            "growth_rate": rate, "time_window": [info0["time"], info1["time"]], "amp": [a0[0], a1[0]],
            "stage_ms": stage_ms, "stage_launches_per_step": n_stage / steps, "fill_ms": fill_ms, "sort_ms": sort_ms,
            "dt_ms": dt_ms}


def growth_fit(t, a):
    """Growth rate of |N_emu| over its exponential phase: the least-squares
    slope of ln a(t) from e^3 above the starting amplitude to e^3 below the
    maximum (before the maximum)."""
    import numpy as np
    t, la = np.asarray(t), np.log(np.asarray(a))
    isat = int(np.argmax(la))
    sel = (np.arange(len(la)) <= isat) & (la >= la[0] + 3.0) & (la <= la.max() - 3.0)
    if sel.sum() < 8:
        return float("nan"), None
    return float(np.polyfit(t[sel], la[sel], 1)[0]), [float(t[sel][0]), float(t[sel][-1])]


def precision_check(core, configs, nsteps=2000):
    """FP32-bulk against the plain FP64 step through saturation on the 1-D
    fiducial workload (1 x 1 x 1024 cells over 64 cm, 378 directions, 387,072
    particles, single mode k = 2 pi/64 cm^-1 plus noise): the growth rate of
    |N_emu| over the exponential phase, the saturation amplitude and time, the
    cell-averaged N_ab in the linear phase and at the end (over the largest
    N_00), and the largest |tr f - 1|.  The FP64 step equals the FP64 oracle to
    1e-12 over the workload's first 20 steps (the GPU tests hold that check);
    2000 oracle steps would take minutes, so it stands in for the oracle."""
    import numpy as np
    parms = configs.fiducial1d()
    emu = core.Emu(parms)
    c0 = core.FlavoredNeutrinoContainer.init_particles(emu)
    K, nh, nf = emu.K, emu.nh, parms.nf
    diag = [K[f"f{i}{i}_Re"] for i in range(nf)]
    res = {}
    for prec in ("fp64", "fp32"):
        c = core.FlavoredNeutrinoContainer(emu, c0.np)
        c.copyParticles(c0)
        st = core.FusedState(emu, c.np, prec)
        st.load(c)
        ts, amps, nab, drift = [], [], {}, 0.0
        info = st.info()
        for n in range(0, nsteps + 1, 10):
            if n:
                st.evolve(10)
                info = st.info()
            ts.append(info["time"])
            amps.append(st.flavor_amplitude()[0])
            if n and n % 200 == 0:
                st.store(c)
                tr = c.rdata[diag].sum(dim=0)
                drift = max(drift, float((tr - 1).abs().max().item()))
                state = core.MultiFab(emu)
                core.deposit_to_mesh(c, state)
                nab[n] = state.valid()[3:3 + nh].reshape(nh, -1).mean(dim=1).cpu().numpy()
        st.close()
        res[prec] = (np.array(ts), np.array(amps), nab, drift)
    (t64, a64, n64, d64), (t32, a32, n32, d32) = res["fp64"], res["fp32"]
    r64, w64 = growth_fit(t64, a64)
    r32, w32 = growth_fit(t32, a32)
    i64, i32 = int(np.argmax(a64)), int(np.argmax(a32))
    lin = max([n for n in n64 if a64[n // 10] < 1e-2 * a64.max()], default=min(n64))
    scale = abs(n64[lin][0])
    return {
        "workload": "fiducial1d: 3 flavours, 1x1x1024 cells over 64 cm, 378 directions/cell, 387072 particles, "
                    f"{nsteps} steps",
        "reference": "plain FP64 HIP step (equal to the FP64 oracle to 1e-12 over 20 steps)",
        "growth_rate_fp64": r64, "growth_rate_fp32": r32,
        "growth_rate_rel_err": abs(r32 - r64) / abs(r64) if r64 == r64 and r64 != 0 else None,
        "growth_window_s": w64, "amplitude_growth": float(a64.max() / a64[0]),
        "saturation_amplitude_rel_diff": float(abs(a32[i32] - a64[i64]) / a64[i64]),
        "saturation_time_rel_diff": float(abs(t32[i32] - t64[i64]) / t64[i64]),
        "cell_avg_Nab_err_linear": float(np.abs(n32[lin] - n64[lin]).max() / scale), "linear_step": lin,
        "cell_avg_Nab_err_end": float(np.abs(n32[nsteps] - n64[nsteps]).max() / scale),
        "trace_drift_fp64": d64, "trace_drift_fp32": d32,
    }


def cpu_baseline(parms_dict, budget_s=15.0):
    """FP64 oracle, OpenMP over this host's cores, on a 16^3-cell box of the
    same physics (same cell size, directions and densities), for up to 10
    steps or the time budget.  For the cpu16 workload (16^3 cells, 10 steps)
    that box is the workload itself, run in full."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(16, ncores))
    d = dict(parms_dict)
    cell = (d["Lx"] / d["ncell"][0], d["Ly"] / d["ncell"][1], d["Lz"] / d["ncell"][2])
    d["ncell"] = (16, 16, 16)
    d["Lx"], d["Ly"], d["Lz"] = 16 * cell[0], 16 * cell[1], 16 * cell[2]
    deck = O.Deck(**{k: v for k, v in d.items() if k in O.Deck.__dataclass_fields__})
    parts = O.init_particles(deck)
    prm = deck.params(nthreads=threads)
    g = O.Grid(deck.nf, prm)
    vv = O.vvac_max(deck.nf, prm, parts)
    O.deposit_to_mesh(deck.nf, prm, parts, g)
    err, dt = O.compute_dt(deck.nf, prm, g, vv)
    steps = 0
    t0 = time.perf_counter()
    while True:
        err, dt, _ = O.rk4_step(deck.nf, prm, parts, g, dt, vv)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 10:
            break
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": len(parts) * steps / el, "unit": "particle-updates/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle (OpenMP, AoS, RK4 Copy/Saxpy) on 16^3 cells x {len(parts) // 4096} directions "
                      f"= {len(parts)} particles, {steps} steps in {el:.1f} s, same cell size and densities; "
                      f"{threads} threads of {ncores} ({cpu_model})"}


def run_adjoint(args):
    """One gradient of J = sum over cells |N_emu|^2 after nsteps RK4 steps with
    respect to the initial flavour state and the antineutrino weight ratio:
    the forward run keeping checkpoints, then the reverse sweep (recompute
    per segment and the VJP of every stage).  Particle-steps/s counts np x
    nsteps for the whole gradient.  Algorithmic bytes: the forward formula
    once for the forward run, once for the recompute and twice for the reverse
    (one scatter and one gather per stage), all FP64."""
    import torch
    from emu_amd import adjoint, configs, core
    workload = "adjoint32" if args.workload == "ff64" else args.workload
    nsteps = args.steps if args.steps != 50 else 200
    parms = configs.WORKLOADS[workload]()
    emu = core.Emu(parms)
    parts = core.FlavoredNeutrinoContainer.init_particles(emu)
    prec = "fp32" if args.precision == "fp32" else "fp64"
    large = parts.np > 50_000_000  # the 128^3 box: no second problem for a warm-up, nothing held twice
    st = core.FusedState(emu, parts.np, "fp64")
    st.load(parts)
    dt = st.info()["dt"]  # compute_dt of the initial state (FP64), held fixed over the run
    st.close()
    del st
    cp = max(1, min(args.checkpoint_every, nsteps))
    if not large:
        warm = adjoint.AdjointProblem(emu, parts, dt=dt, nsteps=min(2, nsteps), checkpoint_every=1, precision=prec)
        warm.gradients(warm.f0.clone(), 1.0)
        del warm
    prob = adjoint.AdjointProblem(emu, parts, dt=dt, nsteps=nsteps, checkpoint_every=cp, segment=args.segment,
                                  precision=prec, snapshots=args.snapshots)
    np_total = parts.np
    f0 = [prob.f0]  # the run takes over f0's memory: nothing else holds it
    prob.f0 = None
    del parts
    footprint = prob.footprint_bytes()
    torch.cuda.empty_cache()
    torch.cuda.reset_peak_memory_stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J, lam, cps, Nn, order = prob.run(f0, 1.0, consume=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    peak_fwd = torch.cuda.max_memory_allocated() / 2**30
    torch.cuda.reset_peak_memory_stats()
    print(f"forward run: {nsteps} steps in {t1 - t0:.1f} s, {peak_fwd:.1f} GiB at the peak", file=sys.stderr,
          flush=True)

    def progress(n):
        # a line every 10 reverse steps on long sweeps (the 128^3 box)
        if large and n % 10 == 0:
            print(f"reverse step {n} done at {time.perf_counter() - t1:.1f} s, "
                  f"{torch.cuda.max_memory_allocated() / 2**30:.1f} GiB at the peak so far", file=sys.stderr,
                  flush=True)

    lam = [lam]  # the sweep holds dJ/dy_N alone, so that it goes after the first reverse step
    g = prob.adjoint_sweep(lam, cps, Nn, order, progress=progress)
    dratio = float(prob.ratio_gradient().item())
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    ncell = parms.ncell[0] * parms.ncell[1] * parms.ncell[2]
    ppc = np_total / ncell
    bps = 4 * algorithmic_bytes(parms.nf, ppc, prec)
    total = t2 - t0
    achieved = np_total * nsteps * bps / total / 1e9
    out = {
        "metric": "adjoint gradient (forward with checkpoints + reverse sweep), particle-steps/s",
        "value": np_total * nsteps / total, "unit": "particle-steps/s", "n_gpus": 1, "steps": nsteps,
        "ms_per_step": total / nsteps * 1e3, "higher_is_better": True, "dtype": "f64" if prec == "fp64" else "f32",
        "data": "synthetic (seeded initial conditions of the workload, generated on the GPU)",
        "config": {"workload": f"{workload}: {parms.nf} flavours, {parms.ncell[0]}x{parms.ncell[1]}x{parms.ncell[2]} "
                               f"periodic cells, {round(ppc)} directions/cell, {np_total} particles",
                   "particles": np_total, "checkpoint_every": cp, "segment": args.segment, "precision": prec,
                   "dt_s": dt, **({"snapshots": prob.snapshots} if args.segment == "binomial" else {})},
        "footprint_gib": footprint / 2**30,
        "peak_allocated_gib": max(peak_fwd, torch.cuda.max_memory_allocated() / 2**30),
        "peak_forward_gib": peak_fwd, "peak_reverse_gib": torch.cuda.max_memory_allocated() / 2**30,
        "forward_ms_per_step": (t1 - t0) / nsteps * 1e3, "reverse_ms_per_step": (t2 - t1) / nsteps * 1e3,
        "replayed_steps": prob.steps_run - nsteps,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes_per_particle_step": bps},
        "J": J, "dJ_dratio": dratio, "grad_f0_norm": float(g.norm().item()),
    }
    print(json.dumps(out), flush=True)


def time_per_function(core, emu, parts, steps, warmup):
    """ms of `steps` steps of the per-function FP64 path (core.Integrator),
    after `warmup` untimed ones, from the given particles (which the
    integrator's double buffering overwrites)."""
    import torch
    integ = core.Integrator(emu, parts)
    integ.integrate(warmup)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    integ.integrate(steps)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1)
    del integ
    torch.cuda.empty_cache()
    return ms


def run_plain(args):
    """The per-function drop-in path (core.Integrator: deposit_to_mesh,
    FillBoundary, copyParticles, interpolate_rhs_from_mesh and the RK4
    Copy/Saxpy calls, then SyncLocation / RedistributeLocal / Renormalize /
    compute_dt) at the workload, FP64: particle-updates/s over the timed
    steps and the fraction of HBM peak the algorithmic bytes give."""
    import torch
    from emu_amd import configs, core
    parms = configs.WORKLOADS[args.workload]()
    emu = core.Emu(parms)
    parts = core.FlavoredNeutrinoContainer.init_particles(emu)
    if args.shuffle:
        # the particles in a random order (the reference's containers hold them in
        # no particular order): the deposit and gather then list them by box
        g = torch.Generator(device="cpu").manual_seed(1)
        perm = torch.randperm(parts.np, generator=g).to(parts.pos.device)
        parts.pos = parts.pos[:, perm].contiguous()
        parts.rdata = parts.rdata[:, perm].contiguous()
        parts.ids = parts.ids[perm].contiguous()
    ms = time_per_function(core, emu, parts, args.steps, args.warmup)
    ncell = parms.ncell[0] * parms.ncell[1] * parms.ncell[2]
    ppc = parts.np / ncell
    bpu = algorithmic_bytes(parms.nf, ppc, "fp64")
    rate = parts.np * args.steps / (ms * 1e-3)
    achieved = rate * bpu / 1e9
    out = {
        "metric": "particle-updates/s of the per-function FP64 drop-in path", "value": rate,
        "unit": "particle-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms / args.steps, "higher_is_better": True, "dtype": "f64",
        "data": "synthetic (seeded initial conditions of the workload, generated on the GPU)",
        "config": {"workload": f"{args.workload}: {parms.nf} flavours, {parms.ncell[0]}x{parms.ncell[1]}x{parms.ncell[2]} "
                               f"periodic cells, {round(ppc)} directions/cell, {parts.np} particles",
                   "particles": parts.np, "path": "plain", "order": "shuffled" if args.shuffle else "cell"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes_per_update": bpu,
                     "note": "wall time of the whole step (every launch), not one kernel"},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="ff64")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-per-function", action="store_true",
                    help="skip timing the per-function FP64 path beside the fused steps")
    ap.add_argument("--no-precision-check", action="store_true",
                    help="skip the FP32-bulk vs FP64 run through saturation on the 1-D fiducial workload")
    ap.add_argument("--precision", default="both", choices=["both", "fp64", "fp32"])
    ap.add_argument("--path", default="fused", choices=["fused", "plain"],
                    help="plain: time the per-function drop-in path (core.Integrator) instead of the fused step")
    ap.add_argument("--adjoint", action="store_true",
                    help="time one gradient through the hand-written adjoint instead (forward with checkpoints and "
                         "the reverse sweep; the 2-flavour 32^3 x 92 workload, 200 steps, checkpoints every 20, "
                         "unless --workload / --steps say otherwise)")
    ap.add_argument("--shuffle", action="store_true",
                    help="--path plain: the particles in a random order instead of the initializer's cell order")
    ap.add_argument("--segment", default="keep", choices=["keep", "store", "recompute", "binomial"],
                    help="--adjoint: what the reverse sweep holds per segment")
    ap.add_argument("--snapshots", type=int, default=2,
                    help="--adjoint --segment binomial: step states held at once per segment")
    ap.add_argument("--checkpoint-every", type=int, default=20, help="--adjoint: steps per checkpoint")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device (5 times): the known-size kernel that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this build runs on a single MI355X; "
              "the multi-GPU decomposition is a design, not an implementation)", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": n * 8, "copies": 5}), flush=True)
        return
    from emu_amd import capi, configs, core
    capi.lib()  # fails loudly without the HIP library
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.adjoint:
        run_adjoint(args)
        return
    if args.path == "plain":
        run_plain(args)
        return
    parms = configs.WORKLOADS[args.workload]()
    emu = core.Emu(parms)
    parts = core.FlavoredNeutrinoContainer.init_particles(emu)
    torch.cuda.synchronize()
    ncell = parms.ncell[0] * parms.ncell[1] * parms.ncell[2]
    ppc = parts.np / ncell
    res = {}
    precs = ["fp64", "fp32"] if args.precision == "both" else [args.precision]
    for prec in precs:
        res[prec] = run_precision(core, parms, parts, prec, args.steps, args.warmup)
        torch.cuda.empty_cache()
    head = "fp32" if "fp32" in res else "fp64"
    r = res[head]
    bpu = algorithmic_bytes(parms.nf, ppc, head)
    # a step is stage_launches_per_step launches of k_stage (four RK stages and the
    # deposit of the step's result); the step's algorithmic bytes are spread over them
    def achieved_gbs(rr, b):
        if rr["stage_ms"] <= 0:
            return None
        return parts.np * b / rr["stage_launches_per_step"] / (rr["stage_ms"] * 1e-3) / 1e9

    achieved = achieved_gbs(r, bpu)
    out = {
        "metric": "particle-updates/s FP32-bulk vs FP64 plain, 1 GPU, at <=1% growth err; %HBM peak",
        "value": r["updates_per_s"],
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head == "fp32" else "f64",
        "data": "synthetic (seeded initial conditions of the workload, generated on the GPU)",
        "config": {"workload": f"{args.workload}: {parms.nf} flavours, {parms.ncell[0]}x{parms.ncell[1]}x{parms.ncell[2]} "
                               f"periodic cells, {round(ppc)} directions/cell, {parts.np} particles, fast-flavor",
                   "particles": parts.np, "precision": head},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": counter_traffic(args.workload, head),
                     "kernel": "k_stage (gather + RHS + RK update + deposit; four RK stages and the step-end deposit)",
                     "algorithmic_bytes_per_update": bpu, "avg_launch_ms": r["stage_ms"],
                     "launches_per_step": r["stage_launches_per_step"]},
    }
    if "fp64" in res:
        f64 = res["fp64"]
        b64 = algorithmic_bytes(parms.nf, ppc, "fp64")
        out["fp64"] = {"value": f64["updates_per_s"], "ms_per_step": f64["ms_per_step"], "stage_ms": f64["stage_ms"],
                       "roofline_frac": (achieved_gbs(f64, b64) or 0.0) / HBM_PEAK_GBS,
                       "growth_rate": f64["growth_rate"]}
    if "fp32" in res and "fp64" in res:
        g32, g64 = res["fp32"]["growth_rate"], res["fp64"]["growth_rate"]
        out["fp32_vs_fp64"] = {"speedup": res["fp32"]["updates_per_s"] / res["fp64"]["updates_per_s"],
                               "growth_rate_rel_err": abs(g32 - g64) / abs(g64) if g64 else None,
                               "growth_rate_fp32": g32, "growth_rate_fp64": g64,
                               "window_s_fp32": res["fp32"]["time_window"], "window_s_fp64": res["fp64"]["time_window"],
                               "amplitude_fp32": res["fp32"]["amp"], "amplitude_fp64": res["fp64"]["amp"]}
    if "fp32" in res and not args.no_per_function:
        # the FP32-bulk step against the per-function FP64 drop-in path (the
        # reference's call structure through the C-ABI) on the same particles
        pf_steps = 3
        ms_pf = time_per_function(core, emu, parts, pf_steps, 1) / pf_steps
        out["fp32_vs_per_function_fp64"] = {
            "ms_per_step_per_function": ms_pf, "steps": pf_steps,
            "speedup": ms_pf / res["fp32"]["ms_per_step"],
            "path": "core.Integrator: deposit_to_mesh, FillBoundary, copyParticles, interpolate_rhs_from_mesh and "
                    "the RK4 Copy/Saxpy calls of main.cpp, FP64"}
    out["kernels_ms"] = {p: {k: res[p][k] for k in ("stage_ms", "fill_ms", "sort_ms", "dt_ms")} for p in res}
    # the same algorithmic bytes over the whole step's wall time (every launch)
    out["roofline"]["wall_frac"] = parts.np * bpu / (r["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS
    if "fp32" in res and "fp64" in res and not args.no_precision_check:
        pc = precision_check(core, configs)
        out["precision_check"] = pc
        # the headline's bar: growth-rate error of FP32-bulk at most 1%
        out["fp32_vs_fp64"]["growth_bar_1pct_met"] = bool(
            pc["growth_rate_rel_err"] is not None and pc["growth_rate_rel_err"] <= 0.01
            and out["fp32_vs_fp64"]["growth_rate_rel_err"] is not None
            and out["fp32_vs_fp64"]["growth_rate_rel_err"] <= 0.01)
    if not args.no_cpu:
        d = {k: getattr(parms, k) for k in parms.__dataclass_fields__ if k != "extra"}
        out["cpu_baseline"] = cpu_baseline(d)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
