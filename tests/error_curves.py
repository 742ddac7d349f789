#!/usr/bin/env python3
"""Error-versus-step curves of the device step against the FP64 model.

For each workload the device advances K steps. At selected steps these errors are reported
for u, v, p and smoke, as relative max-abs (max |a - b| / max |b|), for the reference's
float32 arithmetic (bit-identical to it) and for PreciseAdvection (back-trace in cell units):

  step    one device step against one FP64 step taken from the identical device state, with
          the FP64 model following the float32 branch decisions (sweep count of the early
          exit, turbulence gate) of the float32 oracle on that same state;
  rollout the device trajectory against an FP64 trajectory started from the same initial
          state and run freely (accumulated divergence of FP32 against FP64).

  python tests/error_curves.py --out curves.txt [--only cavity4096]
                                                      (needs the MI355X and the oracle build)
"""
import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

FIELDS = ("U", "V", "p", "M")


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def workloads():
    """(key, title, scene builder, steps, marks, free FP64 rollout)."""
    from fluid_amd import configs as C
    return [
        ("jet256", "jet 256x256, SL, 40 sweeps", lambda mk: C.jet_scene(mk, 256, 256, 40), 200,
         (1, 2, 5, 10, 20, 50, 100, 150, 200), True),
        ("karman", "karman 2048x1024, BFECC, confinement 0.3, 100 sweeps", lambda mk: C.karman_scene(mk), 10,
         (1, 2, 5, 10), True),
        ("cavity512", "cavity 512x512, BFECC, 200 sweeps", lambda mk: C.cavity_scene(mk, 512, 512, 200), 16,
         (1, 2, 4, 8, 16), True),
        ("cavity4096", "cavity 4096x4096, BFECC, 200 sweeps (step error only)",
         lambda mk: C.cavity_scene(mk, 4096, 4096, 200), 100, (1, 10, 50, 100), False),
    ]


def fmt(errs):
    return " ".join("%.1e" % e for e in errs) if errs else "-"


def run(out, threads, only):
    from fluid_amd import Fluid
    from fluid_amd import configs as C
    from oracle.pyoracle import OracleFluid
    lines = ["# error vs step of the device step against the FP64 model (relative max-abs, u v p smoke)",
             "# command: python tests/error_curves.py --out <file>" + (" --only " + ",".join(only) if only else ""),
             "# step: one step from the identical state (FP64 follows the float32 sweep count and turbulence gate);",
             "# rollout: free FP64 trajectory from the same initial state;",
             "# ref: the reference's float32 arithmetic (bit-identical to it); precise: PreciseAdvection"]
    for key, name, build, K, marks, free in workloads():
        if only and key not in only:
            continue
        t0 = time.time()
        g = build(lambda *a: Fluid(*a))
        one = build(lambda *a: Fluid(*a))  # one precise step from g's state
        one.PreciseAdvection = True
        if free:
            gp = build(lambda *a: Fluid(*a))
            gp.PreciseAdvection = True
            r64 = build(lambda *a: OracleFluid(*a, prec=64, threads=threads))
        lines.append("## " + name)
        lines.append("%5s  %-30s %-30s %-30s %-30s" % ("step", "step error ref", "step error precise",
                                                     "rollout error ref", "rollout error precise"))
        print(lines[-2], flush=True)
        for s in range(1, K + 1):
            if s in marks:
                st = g.state()
                print("# step %d: float32 oracle step" % s, flush=True)
                o32 = build(lambda *a: OracleFluid(*a, threads=threads))
                o32.load_state(st)
                o32.update(C.DT)
                print("# step %d: FP64 step (%d sweeps)" % (s, o32.last_sweeps()), flush=True)
                o64 = build(lambda *a: OracleFluid(*a, prec=64, threads=threads))
                o64.load_state(st)
                o64.set_forced_sweeps(o32.last_sweeps())
                o64.set_turb_gate(o32.turb_gate())
                o64.update(C.DT)
                del o32
                one.load_state(st)
                one.update(C.DT)
            g.update(C.DT)
            if free:
                gp.update(C.DT)
                r64.update(C.DT)
            if s in marks:
                se = [rel(g.get(k), o64.get(k)) for k in FIELDS]
                sp = [rel(one.get(k), o64.get(k)) for k in FIELDS]
                re = [rel(g.get(k), r64.get(k)) for k in FIELDS] if free else None
                rp = [rel(gp.get(k), r64.get(k)) for k in FIELDS] if free else None
                del o64
                lines.append("%5d  %-30s %-30s %-30s %-30s" % (s, fmt(se), fmt(sp), fmt(re), fmt(rp)))
                print(lines[-1], flush=True)
        lines.append("# %.0f s" % (time.time() - t0))
    with open(out, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="error_curves.txt")
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--only", default="", help="comma-separated workload keys: jet256,karman,cavity512,cavity4096")
    a = ap.parse_args()
    run(a.out, a.threads, [k for k in a.only.split(",") if k])


if __name__ == "__main__":
    main()
