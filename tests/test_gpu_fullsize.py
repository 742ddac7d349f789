"""The full-size workloads in the GPU suite, each a bounded check against the oracle:

- the differentiable rollout at its specified size (512^2 jet, BFECC, confinement 0.01, 80
  sweeps, 64 steps, checkpoint every 8): the device gradient against the oracle's float32
  adjoint along the identical float32 trajectory;
- the batched ensemble with all 256 members for 3 steps, an impulse step among them: batched
  equals solo stepping bit for bit, and the oracle on the three largest members;
- the 4096^2 cavity (BFECC, 200 sweeps) with PreciseAdvection: one device step from its
  step-10 state within 1e-5 (relative max-abs) of one FP64 step from the same state.
"""
import time

import numpy as np
import pytest

from fluid_amd import Fluid, FluidBatch, Rollout
from fluid_amd import configs as C
from oracle.pyoracle import OracleFluid

pytestmark = pytest.mark.gpu

FIELDS = ("U", "V", "newU", "newV", "p", "S", "M", "newM")


def bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


def test_differentiable_rollout_full_size_matches_float32_adjoint():
    t0 = time.time()
    print("rollout 512^2, 64 steps, checkpoint 8: device", flush=True)
    f = C.diff_scene(lambda *a: Fluid(*a), 512, 512, 80)
    f.Confinement = 0.01
    T = C.gaussian_blobs(f.NumX, f.NumY)
    z = np.zeros((f.NumX, f.NumY), np.float32)
    r = Rollout(f, 64, checkpoint_every=8)
    loss = r.forward(C.DT, z, z, T)
    gu, gv = r.backward()
    r.close()
    f.close()
    print("rollout: float32 oracle adjoint (%.0f s so far)" % (time.time() - t0), flush=True)
    o = C.diff_scene(lambda *a: OracleFluid(*a, threads=8), 512, 512, 80)
    o.Confinement = 0.01
    lo, ou, ov = o.rollout(64, C.DT, z, z, T)
    g = np.concatenate([gu.ravel(), gv.ravel()]).astype(np.float64)
    w = np.concatenate([ou.ravel(), ov.ravel()]).astype(np.float64)
    rel = np.linalg.norm(g - w) / np.linalg.norm(w)
    print("rollout: loss %.9g / %.9g, |g| %.3g, rel %.2e (%.0f s)" % (loss, lo, np.linalg.norm(w), rel,
                                                                      time.time() - t0), flush=True)
    assert abs(loss - lo) <= 1e-12 * abs(lo)
    assert np.isfinite(g).all() and rel <= 1e-5


def test_ensemble_all_members_batched_equals_solo_and_oracle():
    specs = C.ensemble_specs(256, seed=4)
    print("ensemble: %d members, %.1f M cells" % (len(specs), sum(s["w"] * s["h"] for s in specs) / 1e6),
          flush=True)
    b = FluidBatch(C.DENSITY, [(s["w"], s["h"]) for s in specs], C.H)
    solo = [Fluid(C.DENSITY, s["w"], s["h"], C.H) for s in specs]
    for m, s, spec in zip(b.members, solo, specs):
        for f in (m, s):
            C.ensemble_member_setup(f, spec, 60)
            C.ensemble_knobs(f, 60)
    C.ensemble_knobs(b, 60)
    big = sorted(range(len(specs)), key=lambda e: -specs[e]["w"] * specs[e]["h"])[:3]
    oracles = {}
    for e in big:
        o = OracleFluid(C.DENSITY, specs[e]["w"], specs[e]["h"], C.H, threads=8)
        C.ensemble_member_setup(o, specs[e], 60)
        C.ensemble_knobs(o, 60)
        oracles[e] = o
    for step in (9, 10, 11):  # step 10 applies the Gaussian impulses
        imp = C.ensemble_impulses(step, specs)
        assert (step == 10) == bool(imp)
        for (e, cx, cy, fx, fy, sg) in imp:
            b.members[e].ApplyForceGaussian(cx, cy, fx, fy, sg)
            solo[e].ApplyForceGaussian(cx, cy, fx, fy, sg)
            if e in oracles:
                oracles[e].ApplyForceGaussian(cx, cy, fx, fy, sg)
        b.update(C.DT)
        for s in solo:
            s.update(C.DT)
        for o in oracles.values():
            o.update(C.DT)
        for e, (m, s) in enumerate(zip(b.members, solo)):
            for k in FIELDS:
                assert bits_equal(m.get(k), s.get(k)), (step, e, k)
        for e, o in oracles.items():
            for k in FIELDS:
                assert bits_equal(b.members[e].get(k), o.get(k)), (step, "oracle", e, k)
        print("ensemble: step %d equal (batched, solo, oracle on members %s)" % (step, big), flush=True)
    b.close()
    for s in solo:
        s.close()


def rel_maxabs(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def test_cavity_4096_precise_step_within_1e5_of_fp64():
    t0 = time.time()
    print("cavity 4096^2 precise: 10 device steps", flush=True)
    g = C.cavity_scene(lambda *a: Fluid(*a))
    g.PreciseAdvection = True
    for _ in range(10):
        g.update(C.DT)
    st = g.state()
    # the turbulence gate |u| > 0.1 (fluid.go:517) of this step, decided on the float32 state
    # after the sources and the solve, so that the FP64 model takes the same branches
    h = C.cavity_scene(lambda *a: Fluid(*a))
    h.load_state(st)
    h.apply_sources()
    h.run_pass("clearP")
    h.run_pass("makeIncompressible", C.DT, 200)
    sweeps = h.stats()["sweeps_run"]
    U, V, S = h.get("U"), h.get("V"), h.get("S")
    gate = (np.sqrt(U * U + V * V) > np.float32(0.1)) & (S > 0)
    gate[0, :] = gate[-1, :] = gate[:, 0] = gate[:, -1] = False
    h.close()
    g.update(C.DT)
    print("cavity 4096^2 precise: FP64 step (%d sweeps; %.0f s so far)" % (sweeps, time.time() - t0), flush=True)
    o = C.cavity_scene(lambda *a: OracleFluid(*a, prec=64, threads=16))
    o.PreciseAdvection = True
    o.load_state(st)
    o.set_forced_sweeps(sweeps)
    o.set_turb_gate(gate.astype(np.uint8).ravel())
    o.update(C.DT)
    errs = {k: rel_maxabs(g.get(k), o.get(k)) for k in ("U", "V", "p", "M")}
    print("cavity 4096^2 precise: step 11 errors %s (%.0f s)" % (errs, time.time() - t0), flush=True)
    assert all(e <= 1e-5 for e in errs.values()), errs
    g.close()
