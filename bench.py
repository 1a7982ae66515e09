#!/usr/bin/env python3
"""Particle-steps/s of the HIP particle step on one MI355X.

Workloads:
  c3 (default): nested grids + Stokes drift.  Parent 600x500x30 (1 km) with a
     3:1 child 300x300x30 over its centre (grid 1 = child), tanh shelf and a
     masked island, 16 drifting eddies + M2 tide, 32 child-only eddies, Stokes
     drift 0.1 exp(2kz) m/s from 30 deg; 4,000,000 particles released half in
     each grid, RK4 + vertical random displacement (idt 120 s, 10 sub-steps of
     12 s), hand-off between the grids every step.
  c2: 400x300x30 single grid of the same family, 1,000,000 particles, with the
     horizontal random walk (Kh 1 m2/s) on.
Both: FP32 fields with FP64 particle state (--precision fp64 for the FP64 form).

A "step" is one internal time step of all particles (one rp_step).  The
timed region covers K steps including the hydro-slot uploads the schedule
asks for; record generation (the analogue of the reference's netCDF reads)
happens before it.  value = active particle-steps / wall seconds.

The FP64 form is run over the same steps with the same counter-based
deviates, giving its rate and the FP32-field position error against it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from romspath_amd import synth  # noqa: E402
from romspath_amd.runner import GpuRun, RecordSource  # noqa: E402
from romspath_amd.schedule import schedule  # noqa: E402

PEAK_HBM_GBS = 8000.0
PEAK_FP64_VALU_TFLOPS = 78.6          # MI355X FP64 vector, FMA counted as 2
PEAK_FP32_MFMA_TFLOPS = 157.3         # MI355X FP32 matrix, dense
# VALU lane-instructions per second: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
PEAK_VALU_LANE_OPS = 256 * 4 * 16 * 2.4e9
# minimal stencil bytes per active particle-step, no reuse between particles:
# per RK stage 88f+12 (h 4f, zeta 12f, u/v/w 72f, masks 12 B) x4, pm/pn 8f,
# post-step depth+zeta 16f + 4 B per grid tested, VTurb (16+12(N+1))f + 4,
# particle state 88 B.  f = field bytes (4 FP32, 8 FP64).
def b_alg(f, N, vturb=True, ngrid=1):
    rk = 4 * (88 * f + 12)
    post = 16 * f + 4 * ngrid
    vt = (16 + 12 * (N + 1)) * f + 4 if vturb else 0
    return rk + 8 * f + post + vt + 88


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# counters of the in-run counter pass, within one pass's limits (4 TCC, 8 SQ):
# L2 hits and misses and 32-byte memory requests (read, write); VALU wave
# instructions, the FP64 FMA/MUL/ADD among them, and matrix instructions
PMC = ["TCC_HIT_sum", "TCC_MISS_sum", "TCC_EA0_RDREQ_DRAM_32B_sum", "TCC_EA0_WRREQ_WRITE_DRAM_32B_sum",
       "SQ_INSTS_VALU", "SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_ADD_F64",
       "SQ_INSTS_MFMA"]
MFMA_FLOP = 4096          # one v_mfma_f32_32x32x2_f32: 32 x 32 x 2 multiply-adds
STEP_KERNELS = ("vwalk_modal_kernel", "vwalk_kernel", "advect_kernel", "step_kernel", "current_fields")


def counter_pass(args, limit_s=360):
    """HBM bytes and L2 hit rate of the step kernels of this command, counted
    in a short child run of the same workload under rocprofv3 --pmc (the
    counters cannot be read from inside a run).  Run before this process
    touches the GPU.  Bytes = 32 x the 32-byte request counters, which read
    exactly 1 GiB on a 1 GiB read-only kernel and on a 1 GiB copy
    (--copy-calibration).  Returns None without rocprofv3 or when the child
    run fails; the caller then falls back to the stored profile."""
    import glob
    import shutil
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3")
    if exe is None:
        log("counter pass: rocprofv3 not found")
        return None
    out = tempfile.mkdtemp(prefix="rp_pmc_", dir="/tmp")
    child = [sys.executable, str(ROOT / "bench.py"), "--steps", "4", "--warmup", "1", "--no-batched",
             "--no-adjoint", "--no-cpu-baseline", "--no-fp64-form", "--no-counters",
             "--workload", args.workload, "--precision", args.precision,
             "--sort-interval", str(args.sort_interval)]
    if args.particles:
        child += ["--particles", str(args.particles)]
    cmd = ["timeout", "-s", "KILL", str(limit_s), exe, "--pmc", *PMC, "--output-format", "csv",
           "-d", out, "-o", "pmc", "--", *child]
    log("counter pass: the same workload, 4 timed steps, under rocprofv3 --pmc " + " ".join(PMC))
    try:
        r = subprocess.run(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"),
                           stdout=subprocess.PIPE, text=True)
        files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
        if r.returncode != 0 or not files:
            log(f"counter pass failed (exit {r.returncode}); using the stored profile")
            return None
        child_line = json.loads(r.stdout.strip().splitlines()[-1])
        got = summarize_counters(files[0], child_line["roofline"]["active_particles_per_launch"])
    except (OSError, ValueError, KeyError, IndexError, ZeroDivisionError) as e:
        log(f"counter pass failed ({e}); using the stored profile")
        return None
    finally:
        shutil.rmtree(out, ignore_errors=True)
    log(f"counter pass: {got['bytes_per_launch'] / 1e9:.2f} GB per step "
        f"({got['bytes_per_particle_step']:.0f} B per particle-step), L2 hit rate {got['l2_hit_rate']:.3f}; "
        f"per particle-step {got['valu_lane_instr_per_particle_step']:.0f} VALU lane-instructions, "
        f"{got['fp64_flop_per_particle_step']:.0f} FP64 FLOP, "
        f"{got['mfma_flop_per_particle_step'] / MFMA_FLOP:.2f} matrix instructions")
    return got


def summarize_counters(path, per_launch):
    """The step kernels' counters of a rocprofv3 counter-collection CSV (one
    row per dispatch and counter), averaged over each kernel's dispatches:
    HBM bytes per dispatch (32 x the 32-byte request counts), the L2 hit rate
    over all of them, and per particle-step (per_launch active particles a
    dispatch) VALU lane-instructions (wave instructions x 64), FP64 FLOP
    (2 FMA + MUL + ADD, x 64) and matrix FLOP."""
    import csv
    tot = {}
    disp = {}
    with open(path, newline="") as fh:
        for row in csv.DictReader(fh):
            kind = next((k for k in STEP_KERNELS if k in row["Kernel_Name"]), None)
            if kind is None:
                continue
            key = (kind, row["Counter_Name"])
            tot[key] = tot.get(key, 0.0) + float(row["Counter_Value"])
            disp.setdefault(kind, set()).add(row["Dispatch_Id"])
    if not disp:
        raise ValueError("no step-kernel dispatches in the counter file")
    by_kernel = {}
    hit = miss = valu = flop = mfma = 0.0
    for k in sorted(disp):
        n = len(disp[k])
        c = {name: tot.get((k, name), 0.0) / n for name in PMC}
        kv = 64.0 * c["SQ_INSTS_VALU"] / per_launch
        kf = 64.0 * (2.0 * c["SQ_INSTS_VALU_FMA_F64"] + c["SQ_INSTS_VALU_MUL_F64"]
                     + c["SQ_INSTS_VALU_ADD_F64"]) / per_launch
        km = c["SQ_INSTS_MFMA"] / per_launch
        by_kernel[k] = {"read_bytes_per_dispatch": 32.0 * c["TCC_EA0_RDREQ_DRAM_32B_sum"],
                        "write_bytes_per_dispatch": 32.0 * c["TCC_EA0_WRREQ_WRITE_DRAM_32B_sum"],
                        "dispatches": n, "valu_lane_instr_per_particle_step": kv,
                        "fp64_flop_per_particle_step": kf, "mfma_per_particle_step": km}
        hit += tot.get((k, "TCC_HIT_sum"), 0.0)
        miss += tot.get((k, "TCC_MISS_sum"), 0.0)
        valu += kv
        flop += kf
        mfma += km
    launch = sum(v["read_bytes_per_dispatch"] + v["write_bytes_per_dispatch"] for v in by_kernel.values())
    return {"bytes_per_launch": launch, "bytes_per_particle_step": launch / per_launch,
            "l2_hit_rate": hit / max(hit + miss, 1.0), "by_kernel": by_kernel,
            "valu_lane_instr_per_particle_step": valu, "fp64_flop_per_particle_step": flop,
            "mfma_flop_per_particle_step": mfma * MFMA_FLOP,
            "child_particles_per_launch": per_launch}


def window_start(prm, steps, warmup):
    """First timed step: at least `warmup` untimed steps, and late enough that
    the timed window holds a hydro record load after the initial ones (the
    first comes at external step 3, ROMSPath.f90:488-493), so uploads are
    always inside the timed region: the window is centred on that load when
    it is shorter than the run-up to it."""
    first_load = 2 * (prm.dt // prm.idt)
    return max(warmup, first_load - steps // 2)


def run_form(cfg, records, steps, warmup, precision, sort_interval=16):
    prm = cfg.params
    start = window_start(prm, steps, warmup)
    sched = list(schedule(prm.dt, prm.idt, prm.days, prm.iprint, start + steps))
    run = GpuRun(cfg, precision=precision, records=records, sort_interval=sort_interval)
    ctx = run.ctx

    def do(st):
        for ld in st.loads:
            for g in range(1, cfg.ngrid + 1):
                ctx.set_hydro_slot(g, ld.slot, records.get(g, ld.record))
        ctx.step(st.ex, st.ix, st.ets, st.tb, st.tc, st.tf, st.step_index, want_diag=st.output)

    for st in sched[:start]:
        do(st)
    ctx.synchronize()
    a0 = ctx.active_steps()
    ctx.kernel_timing()
    uploads = sum(len(st.loads) for st in sched[start:]) * cfg.ngrid
    t0 = time.perf_counter()
    for k, st in enumerate(sched[start:]):
        do(st)
        if (k + 1) % 60 == 0:
            log(f"  {precision}: {k + 1}/{steps} steps")
    ctx.synchronize()
    t1 = time.perf_counter()
    a1 = ctx.active_steps()
    avg_ms, launches = ctx.kernel_timing()
    phase_ms = ctx.phase_timing()
    capped = ctx.secant_capped()
    out = ctx.get_particles()
    run.close()
    return dict(active=a1 - a0, seconds=t1 - t0, avg_kernel_ms=avg_ms, launches=launches,
                phase_ms=phase_ms, capped=capped, state=out, window=(start, start + steps),
                uploads=uploads)


def batched_ensembles(runup, window, tune_steps, log_every=120):
    """Release ensembles (C4): the same window of steps once as one HIP-graph
    batch per external step over all scenarios, and once as one plain launch
    sequence per scenario, from the same state.  Returns a dict for the line."""
    from romspath_amd.autotune import autotune
    cfg, counts = synth.config4()
    prm = cfg.params
    n = cfg.particles.n
    log(f"C4: {len(counts)} scenarios, {n} particles, run-up {runup} steps, window {window}")
    steps = list(schedule(prm.dt, prm.idt, prm.days, prm.iprint, runup + window))
    records = RecordSource(cfg)
    for st in steps:
        for ld in st.loads:
            records.get(1, ld.record)
    run = GpuRun(cfg, precision="fp32", records=records, sort_interval=16)
    ctx = run.ctx

    def groups(lo, hi):
        i = lo
        while i < hi:
            j = i + 1
            while j < hi and not steps[j].loads:
                j += 1
            yield i, j
            i = j

    # run-up in graph batches; the autotuner replays the last tune_steps
    # steps of the run-up from a snapshot once per candidate, then the
    # run-up's last steps run with the chosen knobs
    tune_from = max(0, runup - tune_steps)
    for i, j in groups(0, tune_from):
        run.load(steps[i])
        ctx.step_batch(steps[i:j], use_graph=True)
        if i // log_every != j // log_every:
            log(f"  run-up {j}/{runup}")
    snap0 = run.snapshot()

    def run_step(st):
        run.load(st)
        ctx.step(st.ex, st.ix, st.ets, st.tb, st.tc, st.tf, st.step_index, want_diag=st.output)

    tuned = autotune(ctx, steps[tune_from:runup], run_step,
                     lambda: run.restore(snap0, steps, tune_from), log=log)
    for st in steps[tune_from:runup]:       # the window again, now with the tuned knobs
        run_step(st)
    snap = ctx.get_particles()
    state = synth.Particles(snap["x"], snap["y"], snap["z"], cfg.particles.dob.copy(),
                            gid=snap["gid"], status=snap["status"], age=snap["age"],
                            isin=snap["isin"])
    b = tuned.best

    def timed(form):
        ctx.set_particles(state)
        ctx.set_tuning(b.block, b.sort_interval, b.tile)
        if form == "plain":
            ctx.set_scenarios(counts)
        ctx.synchronize()
        a0 = ctx.active_steps()
        t0 = time.perf_counter()
        for i, j in groups(runup, runup + window):
            run.load(steps[i])
            if form == "batched":
                ctx.step_batch(steps[i:j], use_graph=True)
            else:
                for st in steps[i:j]:
                    for s_ in range(len(counts)):
                        ctx.step_scenario(s_, st.ex, st.ix, st.ets, st.tb, st.tc, st.tf,
                                          st.step_index, want_diag=st.output)
        ctx.synchronize()
        t1 = time.perf_counter()
        act = ctx.active_steps() - a0
        out = ctx.get_particles()
        log(f"  {form}: {(t1 - t0) / window * 1e3:.2f} ms/step, {act / (t1 - t0):.4g} particle-steps/s")
        return t1 - t0, act, out

    tb_, act_b, out_b = timed("batched")
    tp_, act_p, out_p = timed("plain")
    run.close()
    same = all(np.array_equal(out_b[k], out_p[k]) for k in ("x", "y", "z", "status", "gid"))
    return {
        "workload": "C4: C2 grid and fields, 256 release scenarios of round(exp(U[ln 1e3, ln 2e5])) "
                    "particles in Gaussian clouds (sigma 2 km) around random wet cells, random depths, "
                    "start times U[0, 24 h); FP32 fields",
        "scenarios": int(len(counts)), "particles": int(n),
        "window_steps": window, "runup_steps": runup,
        "active_particle_steps": int(act_b),
        "batched": {"ms_per_step": tb_ / window * 1e3, "value": act_b / tb_,
                    "form": "one HIP graph per external step (30 internal steps), all scenarios"},
        "plain": {"ms_per_step": tp_ / window * 1e3, "value": act_p / tp_,
                  "form": "rp_step_scenario per scenario per step (3 launches each)"},
        "speedup": tp_ / tb_,
        "identical_results": bool(same),
        "tuned": tuned.as_dict(),
    }


def adjoint_cost(n, steps, ckpt, fd_check=True):
    """Adjoint (C5): C2 grid and fields, n particles, vertical random walk on,
    16x16 velocity perturbation (zero), checkpoints every ckpt steps, over a
    window of `steps` internal steps.  Loss = sum of squared horizontal
    distances (km) from final positions to seeded targets minus a soft count
    (box of 20 x 20 km smoothed by a Gaussian of 2 km) of particles inside a
    seeded target box.  Times: the recorded forward (the step the adjoint
    differentiates), the reverse sweep, and the plain FP32-field forward."""
    from scipy.special import ndtr
    base = synth.config2(n_particles=n, seed=5, days=2.0)
    cfg = base
    prm = cfg.params
    g = cfg.grids[0]
    records = RecordSource(cfg)
    for st in schedule(prm.dt, prm.idt, prm.days, prm.iprint, steps):
        for ld in st.loads:
            records.get(1, ld.record)
    rng = np.random.default_rng(55)
    tx = rng.uniform(5.0, g.L - 5.0, n)
    ty = rng.uniform(5.0, g.M - 5.0, n)
    bx0, by0 = rng.uniform(20.0, g.L - 40.0), rng.uniform(20.0, g.M - 40.0)
    sig, beta = 2.0, 10.0

    def soft(u, lo):
        a = (u - lo) / sig
        b = (u - lo - 20.0) / sig
        phi = ndtr(a) - ndtr(b)
        dphi = (np.exp(-0.5 * a * a) - np.exp(-0.5 * b * b)) / (sig * np.sqrt(2 * np.pi))
        return phi, dphi

    def loss_terms(o):
        dx, dy = o["x"] - tx, o["y"] - ty
        px, dpx = soft(o["x"], bx0)
        py, dpy = soft(o["y"], by0)
        L = dx * dx + dy * dy - beta * px * py
        lam = (2 * dx - beta * dpx * py, 2 * dy - beta * px * dpy, np.zeros_like(dx))
        return L, lam

    def forward(record, parts=None):
        c2 = cfg if parts is None else synth.Config(cfg.name, cfg.grids, cfg.hydro, cfg.params,
                                                   parts, cfg.refine)
        run = GpuRun(c2, precision="fp32", records=records, sort_interval=0)
        run.ctx.set_perturbation(np.zeros((2, 16, 16)))
        if record:
            run.ctx.adjoint_record(ckpt)
        run.ctx.synchronize()
        t0 = time.perf_counter()
        out = run.run(max_steps=steps)
        return run, out, time.perf_counter() - t0

    log(f"C5: adjoint, {n} particles, {steps} steps, checkpoints every {ckpt}")
    run, _, t_plain = forward(False)
    run.close()
    run, out, t_fwd = forward(True)
    L, lam = loss_terms(out)
    t0 = time.perf_counter()
    (gx, gy, gz), gp = run.ctx.adjoint(lam, want_pert=True, pert_shape=(2, 16, 16))
    t_adj = time.perf_counter() - t0
    run.close()
    # release (x, y, s): z0 = s h(x0, y0) with zeta 0 at release
    P = cfg.particles
    I = np.floor(P.x).astype(int)
    J = np.floor(P.y).astype(int)
    fx, fy = P.x - I, P.y - J
    h = g.h
    h00, h10, h11, h01 = h[J - 1, I - 1], h[J - 1, I], h[J, I], h[J, I - 1]
    hh = (1 - fx) * (1 - fy) * h00 + fx * (1 - fy) * h10 + fx * fy * h11 + (1 - fx) * fy * h01
    dhdx = (1 - fy) * (h10 - h00) + fy * (h11 - h01)
    dhdy = (1 - fx) * (h01 - h00) + fx * (h11 - h10)
    s_rel = P.z / hh
    g_s = gz * hh
    g_x = gx + gz * s_rel * dhdx
    g_y = gy + gz * s_rel * dhdy
    log(f"  plain {t_plain:.2f} s, recorded forward {t_fwd:.2f} s, adjoint {t_adj:.2f} s "
        f"({t_adj / t_fwd:.2f} x forward)")
    res = {
        "workload": f"C5: C2 grid and fields, {n} particles, RK4 + vertical random walk, "
                    f"{steps}-step window of the 2-day run, checkpoint every {ckpt} steps, "
                    "FP32 fields, FP64 checkpoints",
        "particles": n, "steps": steps, "checkpoint_every": ckpt,
        "t_forward_recorded_s": t_fwd, "t_adjoint_s": t_adj, "t_forward_fp32_form_s": t_plain,
        "adjoint_over_forward": t_adj / t_fwd,
        "adjoint_over_fp32_form_forward": t_adj / t_plain,
        "loss": float(L.sum()),
        "grad_norms": {"release_x": float(np.linalg.norm(g_x)), "release_y": float(np.linalg.norm(g_y)),
                       "release_s": float(np.linalg.norm(g_s)), "perturbation": float(np.linalg.norm(gp))},
    }
    if fd_check:
        # central differences of the per-particle loss in release x, all particles at once
        fds = []
        ok = np.ones(n, dtype=bool)
        for eps in (1e-4, 1e-5):
            vals = []
            for sgn in (1.0, -1.0):
                parts = synth.Particles(P.x + sgn * eps, P.y.copy(), P.z.copy(), P.dob.copy())
                r, o, _ = forward(False, parts)
                r.close()
                ok &= o["status"] == out["status"]
                vals.append(loss_terms(o)[0])
            fds.append((vals[0] - vals[1]) / (2 * eps))
        floor = 1e-3 * np.median(np.abs(fds[0][ok]))
        trusted = ok & (np.abs(fds[0] - fds[1]) <= 1e-5 * np.maximum(np.abs(fds[0]), floor))
        nrm = float(np.linalg.norm((gx - fds[1])[trusted]) / np.linalg.norm(fds[1][trusted]))
        res["fd_check_release_x"] = {"eps": [1e-4, 1e-5], "trusted_particles": int(trusted.sum()),
                                     "particles": n, "normwise_relative_error": nrm}
        log(f"  FD check d/dx0: {trusted.sum()}/{n} trusted, norm-wise rel {nrm:.2e}")
    return res


def cpu_baseline(cfg, n_sample, steps, threads=1, records=None):
    """The oracle on the first n_sample particles for `steps` steps.  Timed:
    the particle-step calls only (hydro records, like the GPU run's, are
    generated outside the timed region; injected deviates are drawn outside)."""
    from oracle import oracle as O
    sub = synth.Particles(cfg.particles.x[:n_sample].copy(), cfg.particles.y[:n_sample].copy(),
                          cfg.particles.z[:n_sample].copy(), cfg.particles.dob[:n_sample].copy())
    scfg = synth.Config(cfg.name, cfg.grids, cfg.hydro, cfg.params, sub, cfg.refine)
    O.set_threads(threads)
    st = O.run_oracle(scfg, max_steps=steps, inject=True, records=records)
    O.set_threads(1)
    dt = st.step_seconds
    return st.active / dt, st.active, dt


def host_cores():
    """CPU threads this process may use (the affinity set, capped by OMP_NUM_THREADS)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(n, 1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=120)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", choices=["c3", "c2"], default="c3")
    ap.add_argument("--particles", type=int, default=None,
                    help="default 4,000,000 (c3) or 1,000,000 (c2)")
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--sort-interval", type=int, default=16,
                    help="re-sort particles by cell every k steps (0 = never)")
    ap.add_argument("--no-fp64-form", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-particles", type=int, default=12000)
    ap.add_argument("--cpu-mt-particles", type=int, default=160000,
                    help="sample of the all-cores oracle run")
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--no-batched", action="store_true", help="skip the C4 ensemble comparison")
    ap.add_argument("--c4-runup", type=int, default=720, help="steps before the C4 window (24 h)")
    ap.add_argument("--c4-window", type=int, default=30)
    ap.add_argument("--c4-tune-steps", type=int, default=64,
                    help="steps each tuning candidate replays (>= 2 x the largest sort interval)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the C5 adjoint measurement")
    ap.add_argument("--no-counters", action="store_true",
                    help="skip the rocprofv3 counter pass (HBM bytes and L2 hit rate of the step)")
    ap.add_argument("--c5-particles", type=int, default=262_144)
    ap.add_argument("--c5-steps", type=int, default=128)
    ap.add_argument("--c5-ckpt", type=int, default=64)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three kernels of known traffic and exit: fill 1 GiB "
                         "(writes 1 GiB), copy 1 GiB device-to-device (reads and writes "
                         "1 GiB), sum 1 GiB (reads 1 GiB): reference readings for the HBM "
                         "byte counters")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        s = a.sum()
        torch.cuda.synchronize()
        log(f"filled, copied and summed 1 GiB (sum {float(s):.0f})")
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path "
              "(the particles would shard by release id, one process per GPU); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    # the counter pass runs first: a child process under the profiler, while
    # this process has not touched the GPU
    counters = None
    if not args.no_counters:
        counters = counter_pass(args)

    if args.workload == "c3":
        npart = args.particles or 4_000_000
        log("building workload C3 (nested 600x500x30 + 3:1 child 300x300x30, Stokes, RK4 + VRW)")
        cfg = synth.config3(n_particles=npart, days=5.0)
        desc = ("C3: nested grids + Stokes drift, parent 600x500x30 + 3:1 child 300x300x30, "
                "16 drifting eddies + M2 tide + 32 child eddies, RK4 + vertical random walk "
                f"(10 x 12 s), {npart} particles half in each grid, hand-off every step")
        dims = {"parent": [600, 500, 30], "child": [300, 300, 30], "refine": 3}
    else:
        npart = args.particles or 1_000_000
        log("building workload C2 (400x300x30, 16 eddies + M2, RK4 + VRW + HRW)")
        cfg = synth.config2(n_particles=npart, days=3.0)
        desc = ("C2: 400x300x30 tanh shelf + island, 16 drifting eddies + M2 tide, "
                "RK4 + vertical random walk (10 x 12 s) + horizontal random walk, "
                f"{npart} particles")
        dims = {"grid": [400, 300, 30]}
    prm = cfg.params
    need = set()
    total = window_start(prm, args.steps, args.warmup) + args.steps
    for st in schedule(prm.dt, prm.idt, prm.days, prm.iprint, total):
        for ld in st.loads:
            need.add(ld.record)
    records = RecordSource(cfg)
    for r in sorted(need):
        for g in range(1, cfg.ngrid + 1):
            records.get(g, r)
    log(f"  {len(need)} hydro records x {cfg.ngrid} grids generated")

    res = run_form(cfg, records, args.steps, args.warmup, args.precision, args.sort_interval)
    value = res["active"] / res["seconds"]
    log(f"  {args.precision}: {value:.4g} particle-steps/s, kernel {res['avg_kernel_ms']:.3f} ms")
    f = 4 if args.precision == "fp32" else 8
    N = cfg.grids[0].N
    balg = b_alg(f, N, vturb=prm.VTurbOn, ngrid=cfg.ngrid)
    per_launch = res["active"] / max(res["launches"], 1)
    achieved = balg * per_launch / (res["avg_kernel_ms"] * 1e-3) / 1e9 if res["avg_kernel_ms"] else 0.0
    adv_ms, walk_ms = res["phase_ms"]
    # HBM bytes and L2 hit rate of this command's step kernels: from the
    # counter pass (a child run under rocprofv3, before this process took the
    # GPU), else from the stored rocprofv3 profile of the same command
    traffic = l2_hit = measured_gbs = None
    tr_src = l2_src = None
    tr_file = ROOT / "profiles" / "hbm_traffic_c3_fp32.json"
    if counters is not None:
        # this run's own counts: bytes per step of the child run's step kernels
        traffic = counters["bytes_per_particle_step"] * per_launch
        l2_hit = counters["l2_hit_rate"]
        if res["avg_kernel_ms"]:
            measured_gbs = traffic / (res["avg_kernel_ms"] * 1e-3) / 1e9
        tr_src = ("counted in this run: 32 x rocprofv3 TCC_EA0_RDREQ_DRAM_32B + "
                  "TCC_EA0_WRREQ_WRITE_DRAM_32B of the step kernels, a child run of the same "
                  "workload (4 timed steps, every dispatch averaged), per particle-step x this "
                  "run's particles per launch")
        l2_src = "counted in this run: TCC_HIT / (TCC_HIT + TCC_MISS) of the step kernels"
    elif args.workload == "c3" and args.precision == "fp32" and tr_file.exists():
        tr_src = ("stored profile of this command: 32 x rocprofv3 TCC_EA0_RDREQ_DRAM_32B + "
                  "TCC_EA0_WRREQ_WRITE_DRAM_32B per particle-step x this run's particles per "
                  "launch, profiles/hbm_traffic_c3_fp32.md")
        l2_src = "stored profile: TCC_HIT / (TCC_HIT + TCC_MISS) of the step kernels, profiles/hbm_traffic_c3_fp32.md"
        tr = json.loads(tr_file.read_text())
        traffic = (tr["read_bytes_per_particle_step"] + tr["write_bytes_per_particle_step"]) * per_launch
        l2_hit = tr.get("l2_hit_rate")
        if res["avg_kernel_ms"]:
            measured_gbs = traffic / (res["avg_kernel_ms"] * 1e-3) / 1e9
    log(f"  kernels: advect {adv_ms:.3f} ms, vertical walk {walk_ms:.3f} ms per step")
    # FP64 vector roofline of the same kernels: FLOP and VALU instructions per
    # particle-step from rocprofv3 SQ counters of this workload
    compute = None
    va_file = ROOT / "profiles" / "valu_c3_fp32.json"
    csrc = None
    if counters is not None and res["avg_kernel_ms"]:
        flop = counters["fp64_flop_per_particle_step"]
        instr = counters["valu_lane_instr_per_particle_step"]
        mfma = counters["mfma_flop_per_particle_step"]
        csrc = ("counted in this run: rocprofv3 SQ_INSTS_VALU(_FMA/_MUL/_ADD_F64), SQ_INSTS_MFMA of "
                "the step kernels in the counter pass's child run, per particle-step")
    elif args.workload == "c3" and args.precision == "fp32" and va_file.exists() and res["avg_kernel_ms"]:
        va = json.loads(va_file.read_text())
        flop = sum(va["fp64_flop_per_particle_step"].values())
        instr = sum(va["valu_lane_instr_per_particle_step"].values())
        mfma = sum(va.get("mfma_per_particle_step", {}).values()) * va.get("mfma_flop", 0)
        csrc = ("stored profile: rocprofv3 SQ_INSTS_VALU(_FMA/_MUL/_ADD_F64), SQ_INSTS_MFMA per "
                "dispatch of this command, profiles/valu_c3_fp32.md")
    if csrc is not None:
        ks = res["avg_kernel_ms"] * 1e-3
        ach = flop * per_launch / ks / 1e12
        mach = mfma * per_launch / ks / 1e12
        compute = {"bound": "fp64 valu", "achieved": ach, "peak": PEAK_FP64_VALU_TFLOPS,
                   "unit": "TFLOP/s", "frac": ach / PEAK_FP64_VALU_TFLOPS,
                   "fp64_flop_per_particle_step": flop,
                   "valu_lane_instr_per_particle_step": instr,
                   "valu_issue_frac": instr * per_launch / ks / PEAK_VALU_LANE_OPS,
                   "fp32_matrix": {"achieved": mach, "peak": PEAK_FP32_MFMA_TFLOPS,
                                   "frac": mach / PEAK_FP32_MFMA_TFLOPS,
                                   "flop_per_particle_step": mfma},
                   "source": csrc}
        if mfma:
            log(f"  FP32 matrix: {mach:.1f} TFLOP/s ({mach / PEAK_FP32_MFMA_TFLOPS:.1%} of peak)")
        log(f"  FP64 VALU: {ach:.1f} TFLOP/s ({ach / PEAK_FP64_VALU_TFLOPS:.1%} of peak), "
            f"VALU issue {compute['valu_issue_frac']:.1%}")

    extra = {}
    if not args.no_fp64_form and args.precision == "fp32":
        r64 = run_form(cfg, records, args.steps, args.warmup, "fp64", args.sort_interval)
        v64 = r64["active"] / r64["seconds"]
        a, b = res["state"], r64["state"]
        same = a["status"] == b["status"]
        same &= a["gid"] == b["gid"]
        layer = np.maximum(b["bath"], 1.0) / N
        ex = np.abs(a["x"] - b["x"])[same]
        ey = np.abs(a["y"] - b["y"])[same]
        ez = (np.abs(a["z"] - b["z"]) / layer)[same]
        exy = np.maximum(ex, ey)
        extra["fp64_form"] = {"value": v64, "avg_kernel_ms": r64["avg_kernel_ms"]}
        extra["fp32_vs_fp64_error_cells"] = {
            "horizontal_max": float(exy.max()), "horizontal_p95": float(np.percentile(exy, 95)),
            "vertical_max": float(ez.max()), "vertical_p95": float(np.percentile(ez, 95)),
            "status_agreement": float(same.mean())}
        log(f"  fp64: {v64:.4g} particle-steps/s")

    if not args.no_batched:
        extra["batched_ensembles"] = batched_ensembles(args.c4_runup, args.c4_window,
                                                       args.c4_tune_steps)

    if not args.no_adjoint:
        extra["adjoint"] = adjoint_cost(args.c5_particles, args.c5_steps, args.c5_ckpt)

    cpu = None
    if not args.no_cpu_baseline:
        log(f"CPU baseline: oracle, {args.cpu_particles} particles x {args.cpu_steps} steps, 1 thread")
        rate1, act1, secs1 = cpu_baseline(cfg, args.cpu_particles, args.cpu_steps, 1, records)
        nt = host_cores()
        log(f"CPU baseline: oracle, {args.cpu_mt_particles} particles x {args.cpu_steps} steps, "
            f"{nt} threads")
        rate, act, secs = cpu_baseline(cfg, args.cpu_mt_particles, args.cpu_steps, nt, records)
        cpu = {"value": rate, "unit": "particle-steps/s", "cores": nt, "kind": "port",
               "sample": f"first {args.cpu_mt_particles} particles of the same workload, "
                         f"{args.cpu_steps} internal steps, FP64 C restatement, OpenMP over "
                         f"particles on {nt} threads, {secs:.1f} s",
               "single_thread": {"value": rate1, "cores": 1,
                                 "sample": f"first {args.cpu_particles} particles, "
                                           f"{args.cpu_steps} steps, {secs1:.1f} s"}}
        log(f"  {rate1:.4g} particle-steps/s on 1 thread, {rate:.4g} on {nt}")

    line = {
        "metric": "particle-steps/s",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["seconds"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("fp32 fields / fp64 state" if f == 4 else "fp64"),
        "data": "synthetic",
        "config": dict({"workload": desc + f", {'FP32' if f == 4 else 'FP64'} fields / FP64 state",
                        "particles": npart, "idt_s": prm.idt, "dt_turb_s": prm.deltat,
                        "field_precision": args.precision, "sort_interval": args.sort_interval,
                        "timed_steps": list(res["window"]),
                        "hydro_uploads_timed": res["uploads"],
                        "parallelism": "single GPU"}, **dims),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "achieved_is": "algorithmic bytes (b_alg x active particles per launch) "
                                    "per second of step kernel time",
                     "traffic_source": tr_src,
                     "traffic_by_kernel": counters["by_kernel"] if counters else None,
                     "l2_hit_rate": l2_hit,
                     "l2_hit_rate_source": l2_src,
                     "measured_traffic_gbs": measured_gbs,
                     "kernel": "one step = current_fields + advect_kernel + vertical walk "
                               "(vwalk_modal_kernel)",
                     "b_alg_bytes_per_particle_step": balg,
                     "avg_kernel_ms": res["avg_kernel_ms"],
                     "advect_kernel_ms": adv_ms, "vwalk_kernel_ms": walk_ms,
                     "active_particles_per_launch": per_launch},
        "compute_roofline": compute,
        "cpu_baseline": cpu,
        "secant_capped": res["capped"],
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
