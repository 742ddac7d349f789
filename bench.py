#!/usr/bin/env python3
"""Throughput of the pkg/fluid simulation step on one MI355X.

Workload (default): the lid-driven cavity at 4096x4096, FP32, seeded solid discs over
>= 5% of the interior, BFECC advection, 200 Gauss-Seidel pressure sweeps per step,
dt 0.1. A "step" is one Update of the scene: per-step lid sources, then Simulate.

Prints one JSON line: cell-steps/s over the timed steps, the roofline of the pressure
solve (algorithmic bytes per solve / its average device time, against 8 TB/s HBM), and
a CPU baseline (the float32 oracle on this host, a bounded sample of the same workload).

Algorithmic bytes per interior cell-step (each pass reads each input field once and
writes each output once, FP32 fields as in the reference, 1-byte cell flags): p clear 4,
GS sweep 28 per sweep, confinement 40, turbulence 20, SL velocity 25, SL smoke 21, BFECC
velocity 67, BFECC smoke 59, viscosity 20.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-steps/s & % HBM peak (FP32, 200 Jacobi) vs Go CPU; adjoint/forward cost"
HBM_PEAK_GBS = 8000.0


def alg_bytes_per_cell(iters, bfecc, conf, turb, visc):
    """Algorithmic HBM bytes per cell of one step: float32 fields, 1-byte cell flags. Advection
    passes count each field they read and write once; the passes that leave newU/newV (newM)
    write them beside U/V (M), both being state the reference keeps. Velocity BFECC: forward 17,
    backward + correction 25, final 25; smoke BFECC: 17, 21, 21; semi-Lagrangian: 25 and 21."""
    b = 4 + 28 * iters
    b += 40 if conf else 0
    b += 20 if turb else 0
    b += 20 if visc else 0
    b += (67 + 59) if bfecc else (25 + 21)
    return b


def build(kind, make, size, iters):
    from fluid_amd import configs as C
    if kind == "cavity":
        return C.cavity_scene(make, size, size, iters)
    if kind == "jet":
        return C.jet_scene(make, size, size, iters)
    if kind == "karman":
        return C.karman_scene(make, 2 * size, size, iters, radius=size // 16, centre=(size // 4, size // 2))
    raise SystemExit("unknown config %s" % kind)


def measured_traffic(kind, size, iters):
    """HBM bytes per GS launch measured with rocprofv3 FETCH_SIZE/WRITE_SIZE for this
    workload (profiles/hbm_traffic_cavity4096.txt), or None when not measured."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_%s%d.txt" % (kind, size))
    if iters != 200 or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("k_gs_wave traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def cpu_baseline(kind, size, iters, threads, early_exit=True):
    """The float32 oracle stepping the same scene on this host's cores: a bounded sample,
    up to three steps (at most ~25 s) of the workload at the largest power-of-two size whose
    step is expected to take under ~20 s (the serial Gauss-Seidel sweeps cost ~15 ns per
    cell-sweep); the median step is reported."""
    from oracle.pyoracle import OracleFluid
    sample_size = size
    while sample_size > 256 and (sample_size ** 2) * iters * 15e-9 > 20.0:
        sample_size //= 2
    f = build(kind, lambda d, w, h, hh: OracleFluid(d, w, h, hh, threads=threads), sample_size, iters)
    f.EarlyExit = early_exit
    times = []
    while len(times) < 3 and sum(times) < 25.0:  # 10-30 s of CPU work, median step
        t0 = time.perf_counter()
        f.update(0.1)
        times.append(time.perf_counter() - t0)
    dt = sorted(times)[len(times) // 2]
    cells = sample_size * sample_size * (2 if kind == "karman" else 1)
    return {"value": cells / dt, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "sample": "float32 oracle (C++ restatement of pkg/fluid, parallelRange threading), median of %d "
                      "%s steps at %dx%d, %.2f s each, %.1f s in all" % (
                          len(times), kind, sample_size * (2 if kind == "karman" else 1), sample_size, dt,
                          sum(times))}


def ensemble_step(target, specs, step, members):
    """One step of the ensemble workload: the step's Gaussian impulses, then Update."""
    from fluid_amd import configs as C
    for (e, cx, cy, fx, fy, sg) in C.ensemble_impulses(step, specs):
        members[e].ApplyForceGaussian(cx, cy, fx, fy, sg)
    for t in target:
        t.update(C.DT)


GO_PUBLISHED_MS = {"quiescent": 1.0, "jet": 4.6}  # Apple M4 Max, fluid_bench_test.go:19,29


def bench_go(args):
    """The reference's own Go benchmarks (fluid_bench_test.go:19-60) on the device: 300x251,
    density 1000, h 0.01, dt 1/120, Simulate's 8 sweeps; "jet" and "bfecc" write U = 4 and
    M = 1 on 100 cells each frame. Frames run back to back as in `go test -bench`; the rate is
    interior cell-steps per second, and vs_baseline divides it by the rate of the frame times
    the reference publishes for an Apple M4 Max (1 ms quiescent, 4.6 ms with the jet)."""
    from fluid_amd import Fluid
    from fluid_amd import configs as C
    v = args.variant
    f = C.go_bench_fluid(lambda *a: Fluid(*a), v)
    f.Solver, f.BlockSweeps, f.HandoffLag = args.solver, args.block_sweeps, args.handoff_lag
    cells = (f.NumX - 2) * (f.NumY - 2)
    for _ in range(args.warmup):
        C.go_bench_frame(f, v)
    f.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        C.go_bench_frame(f, v)
    f.sync()
    wall = time.perf_counter() - t0
    value = cells * args.steps / wall
    pub = GO_PUBLISHED_MS.get(v)
    out = {
        "metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": value / (cells / (pub * 1e-3)) if pub else None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "reference Go benchmark %s: 300x251, dt 1/120, 8 sweeps%s" % (
            {"quiescent": "BenchmarkSimulate", "jet": "BenchmarkSimulateWithJet",
             "bfecc": "BenchmarkSimulateBFECC"}[v], "" if v == "quiescent" else ", jet u=4 on 100 cells"),
            "grid": [300, 251], "published_ms_per_frame": pub},
        "roofline": None, "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        from oracle.pyoracle import OracleFluid
        o = C.go_bench_fluid(lambda d, w, h, hh: OracleFluid(d, w, h, hh, threads=args.cpu_threads), v)
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 10.0:
            C.go_bench_frame(o, v)
            n += 1
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": cells * n / tc, "unit": "cell-steps/s", "cores": args.cpu_threads,
                               "kind": "port", "sample": "float32 oracle, %d frames of the same benchmark "
                                                        "back to back, %.1f s" % (n, tc)}
    print(json.dumps(out), flush=True)


def bench_frame(args):
    """The per-frame outputs around the step, on the cavity state after a few steps: field
    ranges (device reduction, two floats back), Game.Draw's image for each view (computed on
    the device, 4 B per cell downloaded), Vorticity() (4 B per cell downloaded), and 2^20
    tracer particles through SampleVelocityBatch and the RK2 advectParticles (particle arrays
    up and down). Wall time per call with the host transfers; the device kernel times are in
    the rocprof summary of the same command."""
    import numpy as np
    from fluid_amd import Fluid
    from fluid_amd import configs as C
    f = build("cavity", lambda *a: Fluid(*a), args.size, args.iters)
    for _ in range(3):
        f.update(C.DT)
    f.sync()
    W, H = f.NumX - 2, f.NumY - 2
    n = 1 << 20
    rng = np.random.default_rng(1)
    px = (rng.uniform(1.5, f.NumX - 1.5, n) * f.H()).astype(np.float32)
    py = (rng.uniform(1.5, f.NumY - 1.5, n) * f.H()).astype(np.float32)
    age = np.zeros(n, np.float32)
    max_age = np.full(n, 1e9, np.float32)
    ops = {
        "field_range_M": lambda: f.field_range("M"),
        "field_range_p": lambda: f.field_range("p"),
        "render_smoke": lambda: f.render(0),
        "render_pressure": lambda: f.render(1),
        "render_speed": lambda: f.render(2),
        "render_vorticity": lambda: f.render(3),
        "Vorticity": lambda: f.Vorticity(),
        "SampleVelocityBatch_1M": lambda: f.SampleVelocityBatch(px, py),
        "advect_particles_1M": lambda: f.advect_particles(0.01, px.copy(), py.copy(), age.copy(), max_age),
    }
    res = {}
    for name, fn in ops.items():
        for _ in range(2):
            fn()
        f.sync()
        reps = max(3, args.steps // 2)
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        f.sync()
        res[name] = 1e3 * (time.perf_counter() - t0) / reps
    frame_ms = res["render_smoke"] + res["advect_particles_1M"]
    out = {
        "metric": "frame outputs around the step: field ranges, rendered views, tracer particles",
        "value": 1e3 / frame_ms, "unit": "frames/s (smoke image + 2^20 tracer particles)", "n_gpus": 1,
        "steps": reps, "warmup": 2, "ms_per_step": frame_ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "cavity %dx%d after 3 steps; wall ms per call including host transfers" % (W, H),
                   "grid": [W, H], "particles": n},
        "ms_per_call": res, "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_ensemble(args):
    print(json.dumps(run_ensemble(args)), flush=True)


def run_ensemble(args):
    """The batched ensemble: 256 members of seeded sizes 64..1024, presets, wall strokes,
    Gaussian impulses every 10 steps, 60 sweeps, SL, stepped by one FluidBatch (one launch
    per pass, one persistent pressure launch, graph replay, tiles chosen by fl_batch_autotune, or
    the library's default tile shape with --no-autotune), against
    the same members stepped as 256 separate Fluids (one set of launches per grid)."""
    from fluid_amd import Fluid, FluidBatch
    from fluid_amd import configs as C
    specs = C.ensemble_specs(args.members, seed=4)
    cells = sum(s["w"] * s["h"] for s in specs)
    print("bench: ensemble of %d members, %.1f M cells, %d sweeps" % (len(specs), cells / 1e6, args.iters),
          file=sys.stderr, flush=True)
    b = FluidBatch(C.DENSITY, [(s["w"], s["h"]) for s in specs], C.H)
    for m, s in zip(b.members, specs):
        C.ensemble_member_setup(m, s, args.iters)
    C.ensemble_knobs(b, args.iters)
    b.update(C.DT)
    b.sync()
    solve_ms = None
    if not getattr(args, "no_autotune", False):
        tk, solve_ms = b.autotune(C.DT)
        print("bench: autotuned tile depth %d (solve %.2f ms)" % (tk, solve_ms), file=sys.stderr, flush=True)
    else:
        tk = b.BlockSweeps
    for w in range(args.warmup):
        ensemble_step([b], specs, 1 + w, b.members)
    b.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ensemble_step([b], specs, 1 + args.warmup + k, b.members)
    b.sync()
    wall = time.perf_counter() - t0
    ms_batch = 1e3 * wall / args.steps
    sweeps = [m.stats()["sweeps_run"] for m in b.members]
    solve_cell_sweeps = sum(s["w"] * s["h"] * k for s, k in zip(specs, sweeps))
    b_alg_solve = 28.0 * solve_cell_sweeps
    step_bytes = sum(alg_bytes_per_cell(k, False, False, True, False) * s["w"] * s["h"] for s, k in zip(specs, sweeps))
    b.close()

    per_grid = None
    if not args.no_per_grid:
        print("bench: per-grid launches (%d separate Fluids)" % len(specs), file=sys.stderr, flush=True)
        solo = [Fluid(C.DENSITY, s["w"], s["h"], C.H) for s in specs]
        for f, s in zip(solo, specs):
            C.ensemble_member_setup(f, s, args.iters)
            f.BlockSweeps = tk
        for f in solo:
            f.update(C.DT)
        for w in range(args.warmup):
            ensemble_step(solo, specs, 1 + w, solo)
        for f in solo:
            f.sync()
        ns = max(1, args.steps // 2)
        t0 = time.perf_counter()
        for k in range(ns):
            ensemble_step(solo, specs, 1 + args.warmup + k, solo)
        for f in solo:
            f.sync()
        ms_grid = 1e3 * (time.perf_counter() - t0) / ns
        per_grid = {"ms_per_step": ms_grid, "value": cells / (ms_grid * 1e-3), "steps": ns}
        for f in solo:
            f.close()

    out = {
        "metric": METRIC,
        "value": cells / (ms_batch * 1e-3),
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_batch,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "ensemble of %d members, sides U{64,96,..,1024} (seed 4), presets jet/cavity/obstacle, "
                               "4 wall strokes each, Gaussian impulses every 10 steps, %d GS sweeps, SL, dt 0.1"
                               % (len(specs), args.iters),
                   "members": len(specs), "cells": cells, "pressure_sweeps": args.iters, "block_sweeps": tk},
        # the batched solve's own time is known only from the tuner's timing; without
        # it the bound is stated for the whole batched step
        "roofline": ({"bound": "hbm", "achieved": b_alg_solve / (solve_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": b_alg_solve / (solve_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                      "kernel": "batched pressure solve (28 B x sum of member cells x sweeps per launch)"}
                     if solve_ms else
                     {"bound": "hbm", "achieved": step_bytes / (ms_batch * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": step_bytes / (ms_batch * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                      "kernel": "whole batched step (algorithmic bytes of every pass)"}),
        "step_achieved_GBs": step_bytes / (ms_batch * 1e-3) / 1e9,
        "step_frac_hbm": step_bytes / (ms_batch * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "per_grid": per_grid,
        "batched_vs_per_grid": (per_grid["ms_per_step"] / ms_batch) if per_grid else None,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        from oracle.pyoracle import OracleFluid
        print("bench: CPU baseline (float32 oracle, %d threads)" % args.cpu_threads, file=sys.stderr, flush=True)
        done, t0 = 0, time.perf_counter()
        stepped = []
        for s in specs:
            o = OracleFluid(C.DENSITY, s["w"], s["h"], C.H, threads=args.cpu_threads)
            C.ensemble_member_setup(o, s, args.iters)
            t1 = time.perf_counter()
            o.update(C.DT)
            done += time.perf_counter() - t1
            stepped.append(s["w"] * s["h"])
            if done > 10.0:
                break
        out["cpu_baseline"] = {"value": sum(stepped) / done, "unit": "cell-steps/s", "cores": args.cpu_threads,
                               "kind": "port", "sample": "float32 oracle, one step of each of the first %d ensemble "
                                                         "members (%.1f M cells), %.2f s" % (
                                                             len(stepped), sum(stepped) / 1e6, done)}
    return out


def bench_diff(args):
    print(json.dumps(run_diff(args)), flush=True)


def run_diff(args):
    """The differentiable rollout: 512^2 jet (16 rows, u = 1, all walls), BFECC, confinement
    0.01 (0.3 makes the step itself diverge on this grid), 80 sweeps, 64 steps, force field F = 0, loss = |M_64 - T|^2 with T six seeded
    Gaussians (sigma 20 cells), checkpoint every 8 steps. Times the forward rollout and the
    backward pass (segment recompute + adjoint) on the device."""
    import numpy as np
    from fluid_amd import Fluid, Rollout
    from fluid_amd import configs as C
    size, steps, every = args.size if args.size != 4096 else 512, 64, 8
    iters = args.iters if args.iters != 200 else 80
    f = C.diff_scene(lambda *a: Fluid(*a), size, size, iters)
    f.Confinement = args.confinement
    f.BlockSweeps = args.block_sweeps
    f.HandoffLag = args.handoff_lag
    shape = (f.NumX, f.NumY)
    T = C.gaussian_blobs(f.NumX, f.NumY)
    z = np.zeros(shape, dtype=np.float32)
    r = Rollout(f, steps, checkpoint_every=every)
    print("bench: differentiable rollout %d^2, %d steps, %d sweeps" % (size, steps, iters), file=sys.stderr, flush=True)
    r.forward(C.DT, z, z, T)
    r.backward()
    reps = max(1, min(args.steps, 5))
    tf = tb = 0.0
    for _ in range(reps):
        t0 = time.perf_counter()
        loss = r.forward(C.DT, z, z, T)
        t1 = time.perf_counter()
        gu, gv = r.backward()
        t2 = time.perf_counter()
        tf += t1 - t0
        tb += t2 - t1
    tf, tb = tf / reps, tb / reps
    cells = size * size
    st = f.stats()
    b_alg = alg_bytes_per_cell(st["sweeps_run"], True, True, True, False)
    out = {
        "metric": METRIC,
        "value": cells * steps / tf,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 1,
        "ms_per_step": 1e3 * tf / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "differentiable jet rollout %dx%d FP32, BFECC, confinement %g, %d GS sweeps, "
                               "%d steps, checkpoint every %d, F = 0" % (size, size, args.confinement, iters, steps,
                                                                         every),
                   "grid": [size, size], "pressure_sweeps": iters, "steps": steps, "checkpoint_every": every},
        "forward_ms": 1e3 * tf,
        "backward_ms": 1e3 * tb,
        "adjoint_over_forward": tb / tf,
        "loss": loss,
        "grad_norm": float(np.sqrt((gu.astype(np.float64) ** 2).sum() + (gv.astype(np.float64) ** 2).sum())),
        "step_achieved_GBs": b_alg * cells / (tf / steps) / 1e9,
        "roofline": None,
        "cpu_baseline": None,
    }
    r.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="cavity", choices=["cavity", "jet", "karman", "ensemble", "diff", "go", "frame"])
    ap.add_argument("--variant", default="jet", choices=["quiescent", "jet", "bfecc"],
                    help="with --config go: which of the reference's Go benchmarks")
    ap.add_argument("--members", type=int, default=256)
    ap.add_argument("--confinement", type=float, default=0.01,
                    help="differentiable rollout: vorticity confinement (0.3 diverges on the 512^2 jet)")
    ap.add_argument("--no-per-grid", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="ensemble: step with the library's default tile shape instead of fl_batch_autotune's")
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--iters", type=int, default=200)
    ap.add_argument("--solver", type=int, default=0)
    ap.add_argument("--block-sweeps", type=int, default=0)
    ap.add_argument("--handoff-lag", type=int, default=0, help="0 auto, 1 on time, 2 one block late")
    ap.add_argument("--precise", action="store_true",
                    help="PreciseAdvection: back-trace in cell units (an addition; default is the reference's arithmetic)")
    ap.add_argument("--multigrid", action="store_true",
                    help="UseMultigrid with 2 levels; --iters is then the number of V-cycles per step")
    ap.add_argument("--no-early-exit", action="store_true",
                    help="run every pressure sweep (the Karman workload always does: see its note)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="default workload: skip the adjoint/forward and batched/per-grid parts of the metric")
    ap.add_argument("--cpu-threads", type=int, default=min(os.cpu_count() or 1, 16))
    ap.add_argument("--calibrate", action="store_true",
                    help="first run a 256 MiB device-to-device copy (reference kernel for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (the multi-GPU decomposition "
              "is a design only and not built)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from fluid_amd import Fluid
    from fluid_amd import configs as C

    if args.config == "diff":
        return bench_diff(args)
    if args.config == "frame":
        return bench_frame(args)
    if args.config == "go":
        if args.steps == 20:  # frames take tens of microseconds: time a few seconds' worth
            args.steps, args.warmup = 2000, max(args.warmup, 50)
        return bench_go(args)
    if args.config == "ensemble":
        if args.iters == 200:
            args.iters = 60
        return bench_ensemble(args)

    if args.calibrate:
        import torch
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    print("bench: building %s %d^2, %d sweeps" % (args.config, args.size, args.iters), file=sys.stderr, flush=True)
    f = build(args.config, lambda *a: Fluid(*a), args.size, args.iters)
    f.Solver = args.solver
    f.BlockSweeps = args.block_sweeps
    f.HandoffLag = args.handoff_lag
    f.PreciseAdvection = args.precise
    if args.no_early_exit or args.config == "karman":
        f.EarlyExit = False
    if args.multigrid:
        f.UseMultigrid, f.MultigridLevels = True, 2
    W, H = f.NumX - 2, f.NumY - 2
    cells = W * H
    for _ in range(args.warmup):
        f.update(C.DT)
    f.sync()
    print("bench: warm-up done, timing %d steps" % args.steps, file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        f.update(C.DT)
    f.sync()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    value = cells * args.steps / wall
    st = f.stats()

    # Per-pass device times from HIP events on the library's stream.
    f.set_timing(True)
    nt = max(1, min(args.steps, 5))
    for _ in range(nt):
        f.update(C.DT)
    tm = f.timing()
    f.set_timing(False)
    solve_ms = tm["pressure_ms"] / tm["steps"]
    sweeps = st["sweeps_run"]
    solve_bytes = 28.0 * sweeps * cells
    achieved = solve_bytes / (solve_ms * 1e-3) / 1e9
    b_alg = alg_bytes_per_cell(sweeps, f.UseBFECC, f.Confinement != 0, f.TurbulenceStrength > 0,
                               f.ViscosityDiffusion > 0)
    step_gbs = b_alg * cells / (ms_per_step * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "%s %dx%d FP32, %s/step, BFECC %s, dt 0.1" % (
            args.config, W, H, ("%d V-cycles (%d fine GS sweeps, 40 coarse sweeps each)" % (args.iters, sweeps))
            if args.multigrid else "%d GS sweeps" % sweeps, "on" if f.UseBFECC else "off"),
            "grid": [W, H], "pressure_sweeps": sweeps,
            "solver": "multigrid" if args.multigrid else ["auto", "hyperplane", "temporal-blocked"][args.solver],
            "advection": "precise (cell-unit back-trace)" if args.precise else "reference arithmetic"},
        "roofline": None if args.multigrid else {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(args.config, args.size, args.iters),
            "kernel": "pressure solve (%d sweeps x 28 B x %d cells per launch)" % (sweeps, cells),
            "note": "achieved counts the algorithmic bytes of a sweep-by-sweep solve; the temporal-blocked "
                    "kernel moves `traffic` bytes of HBM per launch (rate in hbm_measured) and is bound by its "
                    "critical path of column steps and tile handoffs, not by HBM"},
        "early_exit": bool(f.EarlyExit),
        "step_alg_bytes_per_cell": b_alg,
        "step_achieved_GBs": step_gbs,
        "step_frac_hbm": step_gbs / HBM_PEAK_GBS,
        "device_ms_per_step": {"total": tm["total_ms"] / tm["steps"], "pressure": solve_ms,
                               "advect": tm["advect_ms"] / tm["steps"], "other": tm["other_ms"] / tm["steps"]},
        "cpu_baseline": None,
    }
    out["invalid_sample"] = bool(f.stats()["invalid_sample"])
    if args.config == "karman":
        out["note"] = ("Karman street from rest, timed over steps %d-%d, with the early exit off so that every step "
                       "runs all its sweeps: the workload as specified (inflow 1, disc r = H/16, BFECC, confinement "
                       "0.3, 100 sweeps) has no steady state. In the reference's float32 arithmetic (which the device "
                       "reproduces bit for bit) max|u| reaches 10 by step 9, 1e13 by step 22 and NaN at step 24 at "
                       "2048x1024; use --steps 8 --warmup 2 to time the steps before it runs away"
                       % (args.warmup + 1, args.warmup + args.steps))
    if out["roofline"]:
        out["latency_roof"] = latency_roof(f, solve_ms)
    traffic = out["roofline"]["traffic"] if out["roofline"] else None
    if traffic:
        # the solve's physical HBM traffic (rocprof FETCH_SIZE + WRITE_SIZE per launch, calibrated
        # against a copy kernel) over its live device time: temporal blocking keeps most of the
        # 28 B per cell-sweep on chip, so this is far below the algorithmic rate above
        phys = traffic / (solve_ms * 1e-3) / 1e9
        out["hbm_measured"] = {"traffic": traffic, "achieved": phys, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": phys / HBM_PEAK_GBS,
                               "source": "profiles/hbm_traffic_%s%d.txt" % (args.config, args.size)}
    f.close()
    if not args.no_cpu_baseline:
        print("bench: CPU baseline (float32 oracle, %d threads)" % args.cpu_threads, file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.config, args.size, args.iters, args.cpu_threads, out["early_exit"])
    if args.config == "cavity" and not args.no_extras:
        out.update(metric_extras(args))
    print(json.dumps(out), flush=True)


def critical_path_steps(L):
    """Column steps on the solve's longest chain of dependent tiles, from its tiling (a dict of
    fl_get_gs_layout). The kernel's wait rule: before running block b a tile stages block b + 1,
    so it needs its G producer (the tile one row-tile up) to have published row s1 + 64 and its
    K producer (the tile one sweep-block back) row s1 + TK, s1 being block b + 1's last step.
    Rows are published in whole blocks (on time: when the next block starts; halfway: BLK/2 steps
    later; one block late: BLK steps later), which puts a consumer 2 + ceil((L - 1) / BLK) blocks
    plus that delay behind its producer, L = 64 along the rows and TK along the sweeps. The last
    tile starts after G - 1 row hops and KB - 1 sweep hops, then runs its own steps."""
    tk, blk, lag = L["sweeps_per_tile"], L["cols_per_block"], L["handoff_lag"]
    delay = {1: 0, 2: blk, 3: blk // 2}[lag]

    def hop(span):
        return (2 + -(-(span - 1) // blk)) * blk + delay
    return (L["row_tiles"] - 1) * hop(64) + (L["sweep_blocks"] - 1) * hop(tk) + L["steps_per_tile"]


def column_step_us(L, dt):
    """Wall time of one column step of a tile running alone (no producer to wait for): a 40-row
    strip 4000 columns long, one tile of the same shape and publication, solved 5 times."""
    import numpy as np
    from fluid_amd import Fluid
    f = Fluid(1000.0, 40, 4000, 0.01)
    rng = np.random.default_rng(7)
    st = {k: (0.3 * rng.standard_normal((42, 4002))).astype(np.float32) for k in ("U", "V", "newU", "newV")}
    S = np.ones((42, 4002), np.float32)
    S[0, :] = S[-1, :] = S[:, 0] = S[:, -1] = 0
    st["S"] = S
    f.load_state(st)
    f.EarlyExit = False
    f.BlockSweeps, f.BlockCols, f.HandoffLag = L["sweeps_per_tile"], L["cols_per_block"], L["handoff_lag"]
    tk = L["sweeps_per_tile"]
    for _ in range(2):
        f.run_pass("makeIncompressible", dt, tk)
    f.sync()
    t0 = time.perf_counter()
    for _ in range(5):
        f.run_pass("makeIncompressible", dt, tk)
    f.sync()
    per_solve = (time.perf_counter() - t0) / 5
    one = f.gs_layout()
    f.close()
    assert one["tiles"] == 1, one
    return 1e6 * per_solve / one["steps_per_tile"]


def latency_roof(f, solve_ms):
    """The temporal-blocked solve's latency bound: its critical path of column steps times the
    measured cost of one column step of a lone tile. `frac` is that bound over the solve's
    measured time (how far the solve is from running at the speed of its longest chain)."""
    from fluid_amd import configs as C
    try:
        L = f.gs_layout()
    except Exception:
        return None
    steps = critical_path_steps(L)
    us = column_step_us(L, C.DT)
    ms = steps * us * 1e-3
    return {"critical_path_steps": steps, "column_step_us": us, "ms": ms, "frac": ms / solve_ms,
            "shape": "%d sweeps x %d columns per tile block, %s publication, %d tiles on %d workgroups" % (
                L["sweeps_per_tile"], L["cols_per_block"], {1: "on-time", 2: "one-block-late", 3: "halfway"}[
                    L["handoff_lag"]], L["tiles"], L["workgroups"]),
            "note": "column_step_us is measured on a single tile (a 40x4000 strip, whose edge waves take "
                    "the general update); the rest of frac is time tiles spend in the pipeline's fill and "
                    "drain, waiting on producers and sharing CUs"}


def metric_extras(args):
    """The metric's other two parts, measured in the same run: the adjoint's cost over the
    forward rollout on the differentiable workload (512^2 jet, 64 steps, checkpoint 8), and
    the batched ensemble (256 uneven members) against the same members stepped one grid per
    launch."""
    ns = argparse.Namespace(size=512, iters=80, confinement=args.confinement, block_sweeps=0, handoff_lag=0,
                            steps=3)
    print("bench: adjoint/forward on the differentiable rollout", file=sys.stderr, flush=True)
    d = run_diff(ns)
    ne = argparse.Namespace(members=256, iters=60, steps=10, warmup=3, no_per_grid=False, no_cpu_baseline=True,
                            cpu_threads=args.cpu_threads)
    print("bench: batched ensemble against per-grid launches", file=sys.stderr, flush=True)
    e = run_ensemble(ne)
    return {
        "adjoint_over_forward": {
            "value": d["adjoint_over_forward"], "forward_ms": d["forward_ms"], "backward_ms": d["backward_ms"],
            "target": "<= 3", "workload": d["config"]["workload"],
            "gradient_check": "equals the float32 oracle's adjoint on the identical trajectory to rel <= 1e-5 (GPU "
                              "test suite); finite differences agree to 1e-3 only over 1-2 steps: the rollout "
                              "expands dL/dF 3-10x per step (about 8e28 at 64 steps) and every usable FD step "
                              "crosses a BFECC clamp or the smoke's max(., 0) within 8 steps"},
        "ensemble_batched_over_per_grid": {
            "value": e["batched_vs_per_grid"], "batched_ms_per_step": e["ms_per_step"],
            "per_grid_ms_per_step": e["per_grid"]["ms_per_step"], "batched_cell_steps_per_s": e["value"],
            "target": ">= 3", "workload": e["config"]["workload"]},
    }


if __name__ == "__main__":
    main()
