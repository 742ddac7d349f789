#!/usr/bin/env python3
"""Full-size check of the device adjoint on the differentiable jet rollout.

Two halves, because the oracle cannot run where the GPU is exercised in a test budget:

  python tests/adjoint_check.py gpu --out grads.npz
      (on an MI355X) runs the rollout (512^2 jet, BFECC, confinement 0.01, 80 sweeps,
      F = 0, target six Gaussians) for horizons of 1, 2, 8 and 64 steps and saves the loss
      and dL/dF of each.
  python tests/adjoint_check.py cpu --grads grads.npz --summary out.txt
      (on the host) compares them with the oracle: its float32 adjoint along the identical
      trajectory (every horizon), its FP64 adjoint, and central finite differences of its
      FP64 forward pass on 32 seeded directions (short horizons).

The rollout is strongly expanding (BFECC and the confinement's normalised grad|curl|):
|dL/dF| grows 3-10x per step, so beyond a few steps no FD step that FP64 resolves stays
clear of the clamp and max(.,0) kinks; the long horizons are therefore checked against
the oracle's adjoint on the same float32 trajectory.
"""
import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

SIZE, ITERS, CONF, EVERY = 512, 80, 0.01, 8
HORIZONS = (1, 2, 8, 64)
FD_HORIZONS = (1, 2, 8)
NDIR, EPS, SEED = 32, 1e-8, 11


def scene(make):
    from fluid_amd import configs as C
    f = C.diff_scene(make, SIZE, SIZE, ITERS)
    f.Confinement = CONF
    return f


def run_gpu(out):
    from fluid_amd import Fluid, Rollout
    from fluid_amd import configs as C
    res = {}
    for H in HORIZONS:
        f = scene(lambda *a: Fluid(*a))
        T = C.gaussian_blobs(f.NumX, f.NumY)
        z = np.zeros((f.NumX, f.NumY), dtype=np.float32)
        r = Rollout(f, H, checkpoint_every=EVERY)
        loss = r.forward(C.DT, z, z, T)
        gu, gv = r.backward()
        r.close()
        f.close()
        res["loss_%d" % H] = np.float64(loss)
        res["gu_%d" % H], res["gv_%d" % H] = gu, gv
        print("gpu horizon %d: loss %.9g |g| %.4g" % (H, loss, np.sqrt((gu.astype(np.float64) ** 2).sum() +
                                                                          (gv.astype(np.float64) ** 2).sum())),
              flush=True)
    np.savez(out, **res)


def _fd_worker(a):
    H, k = a
    from oracle.pyoracle import OracleFluid
    from fluid_amd import configs as C
    o = scene(lambda *x: OracleFluid(*x, prec=64))
    shape = (o.NumX, o.NumY)
    T = C.gaussian_blobs(o.NumX, o.NumY).astype(np.float64)
    rng = np.random.default_rng(SEED + 1000 * H + k)
    du, dv = rng.standard_normal(shape), rng.standard_normal(shape)
    lp = o.rollout(H, C.DT, EPS * du, EPS * dv, T, grad=False)
    lm = o.rollout(H, C.DT, -EPS * du, -EPS * dv, T, grad=False)
    return H, k, (lp - lm) / (2 * EPS)


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)  # float32 norms overflow here
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def run_cpu(grads, summary, workers):
    from multiprocessing import Pool
    from oracle.pyoracle import OracleFluid
    from fluid_amd import configs as C
    d = np.load(grads)
    lines = ["# device adjoint vs oracle, differentiable jet rollout %d^2, BFECC, confinement %g, %d sweeps, F = 0"
             % (SIZE, CONF, ITERS),
             "# command: python tests/adjoint_check.py gpu --out G.npz (MI355X); "
             "python tests/adjoint_check.py cpu --grads G.npz --summary S (host)",
             "# rel = |g_device - g_ref|_2 / |g_ref|_2; fd: |<g_device,d> - FD(d)| / |FD(d)| over %d directions, "
             "eps %g, FP64 forward" % (NDIR, EPS)]
    t0 = time.time()
    jobs = [(H, k) for H in FD_HORIZONS for k in range(NDIR)]
    g64 = {}
    with Pool(workers) as pool:
        fd_async = pool.map_async(_fd_worker, jobs)
        for H in HORIZONS:
            G = np.stack([d["gu_%d" % H], d["gv_%d" % H]]).astype(np.float64)
            o32 = scene(lambda *x: OracleFluid(*x, threads=workers))
            T = C.gaussian_blobs(o32.NumX, o32.NumY)
            z = np.zeros((o32.NumX, o32.NumY), dtype=np.float32)
            l32, u32, v32 = o32.rollout(H, C.DT, z, z, T)
            line = "horizon %2d: |g| %.4g  loss device %.12g oracle-f32 %.12g  rel(f32 adjoint, same trajectory) %.2e" % (
                H, np.linalg.norm(G.ravel()), float(d["loss_%d" % H]), l32, rel(G, np.stack([u32, v32])))
            if H <= 8:
                o64 = scene(lambda *x: OracleFluid(*x, prec=64, threads=workers))
                T64 = T.astype(np.float64)
                z64 = np.zeros(T64.shape)
                l64, u64, v64 = o64.rollout(H, C.DT, z64, z64, T64)
                g64[H] = np.stack([u64, v64])
                line += "  rel(FP64 adjoint) %.2e" % rel(G, g64[H])
            lines.append(line)
            print(line, flush=True)
        fd = fd_async.get()
    for H in FD_HORIZONS:
        G = np.stack([d["gu_%d" % H], d["gv_%d" % H]]).astype(np.float64)
        errs, errs64 = [], []
        for (h, k, val) in fd:
            if h != H:
                continue
            rng = np.random.default_rng(SEED + 1000 * H + k)
            shape = G.shape[1:]
            du, dv = rng.standard_normal(shape), rng.standard_normal(shape)
            ad = float((G[0] * du).sum() + (G[1] * dv).sum())
            ad64 = float((g64[H][0] * du).sum() + (g64[H][1] * dv).sum())
            errs.append(abs(ad - val) / max(abs(val), 1e-300))
            errs64.append(abs(ad64 - val) / max(abs(val), 1e-300))
        errs, errs64 = np.array(errs), np.array(errs64)
        line = ("horizon %2d: fd over %d directions: device median %.2e max %.2e (<=1e-3 in %d of %d); "
                "FP64 adjoint median %.2e max %.2e (<=1e-3 in %d)" % (
                    H, len(errs), np.median(errs), errs.max(), int((errs <= 1e-3).sum()), len(errs),
                    np.median(errs64), errs64.max(), int((errs64 <= 1e-3).sum())))
        lines.append(line)
        print(line, flush=True)
    lines.append("# host time %.0f s with %d worker processes" % (time.time() - t0, workers))
    with open(summary, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["gpu", "cpu"])
    ap.add_argument("--out", default="adjoint_grads.npz")
    ap.add_argument("--grads", default="adjoint_grads.npz")
    ap.add_argument("--summary", default="adjoint_check.txt")
    ap.add_argument("--workers", type=int, default=min(8, os.cpu_count() or 1))
    a = ap.parse_args()
    if a.mode == "gpu":
        run_gpu(a.out)
    else:
        run_cpu(a.grads, a.summary, a.workers)


if __name__ == "__main__":
    main()
