"""Drive the host runtime of libfluidhip through the Python mirror: contexts of several sizes,
the edit and scene API, every solver path, views, rendering, particles, staged copies,
batches, the rollout adjoint, the tuners and the error paths. Run by tests/test_host_asan.py
in a child process against the AddressSanitizer build (FLUID_AMD_LIB, the ASan runtime
preloaded), where any heap error in host code aborts the process with a report.
It needs a GPU; it checks only that every call completes (results are checked elsewhere)."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from fluid_amd import Fluid, FluidBatch, Rollout  # noqa: E402
from fluid_amd import _lib  # noqa: E402
from fluid_amd import configs as C  # noqa: E402


def step(f, n, dt=C.DT):
    for _ in range(n):
        f.update(dt)


def main():
    assert os.path.realpath(_lib.LIB_PATH).endswith("libfluidhip_asan.so"), _lib.LIB_PATH
    rng = np.random.default_rng(1)

    # one-cell, small and ragged grids through the default solve
    for (w, h) in ((1, 1), (37, 23), (150, 90)):
        f = Fluid(1000.0, w, h, 0.01)
        st = {k: (0.3 * rng.standard_normal((w + 2, h + 2))).astype(np.float32) for k in ("U", "V", "newU", "newV")}
        st["S"] = (rng.random((w + 2, h + 2)) > 0.1).astype(np.float32)
        f.load_state(st)
        f.PressureIters = 12
        f.Simulate(0.1)
        f.state()
        f.close()

    # the jet with the edit API, BFECC, confinement, turbulence and every view
    f = C.jet_scene(lambda *a: Fluid(*a), 96, 80, 20)
    f.UseBFECC, f.Confinement, f.TurbulenceStrength = True, 0.3, 0.5
    f.SetCircularObstacle(40, 40, 8)
    f.SetSolid(3, 4, True)
    f.SetVelocity(4, 4, 2.5, -1.5)
    f.AddSmoke(5, 5, 0.25)
    f.ApplyForce(6, 6, 1.0, 1.0)
    f.ApplyForceRadius(20, 20, 3.0, -2.0, 4)
    f.ApplyForceGaussian(30, 30, 1.0, 0.5, 3.0)
    f.draw_line(1, 1, 60, 50, True)
    f.smoke_brush(10, 60, 3, 2.0)
    f.add_source(10, 10, 1.0, 0.0)
    f.add_sink(80, 70, 0.0, 0.0)
    step(f, 4)
    f.clear_sources()
    for name in ("Vorticity", "VelocityMagnitude", "Pressure", "Smoke", "Velocity"):
        getattr(f, name)()
    f.MaxDivergence()
    f.GetAdaptiveTimeStep(0.1)
    f.SampleVelocity(0.2, 0.15)
    xs = (rng.random(300) * 0.9).astype(np.float32)
    f.SampleVelocityBatch(xs, xs[::-1].copy())
    for view in range(4):
        f.render(view, overlay=bool(view & 1))
    n = 2000
    px = (rng.random(n) * 0.9).astype(np.float32)
    py = (rng.random(n) * 0.7).astype(np.float32)
    age = np.zeros(n, np.float32)
    mx = np.full(n, 1.0, np.float32)
    f.advect_particles(np.float32(C.DT), px, py, age, mx)
    f.advect_particles(np.float32(C.DT), px, py, age, mx, no_alloc=True)
    f.set_timing(True)
    step(f, 2)
    f.timing()
    f.set_timing(False)
    f.stats()
    f.sweep_maxdiv()
    for bad in ((96 + 2, 0), (0, -1)):
        try:
            f.SetSolid(bad[0], bad[1], True)
        except IndexError:
            pass
        else:
            raise AssertionError("out-of-range index accepted")
    f.last_error()
    for preset in ("jet", "cavity", "karman"):
        f.reset_to_preset(preset, 3.0)
        step(f, 1)
    f.Reset()
    f.close()

    # solver variants: early exit off, multigrid, precise advection, every tile depth, the tuner
    f = C.cavity_scene(lambda *a: Fluid(*a), 200, 140, 40)
    step(f, 2)
    f.EarlyExit = False
    step(f, 1)
    f.UseMultigrid, f.MultigridLevels = True, 2
    step(f, 2)
    f.UseMultigrid = False
    f.PreciseAdvection = True
    step(f, 1)
    f.PreciseAdvection = False
    for tk in (2, 4, 8, 16):
        for blk in (4, 8, 16):
            f.BlockSweeps, f.BlockCols, f.HandoffLag = tk, blk, 1 + (tk + blk) % 3
            step(f, 1)
    f.BlockSweeps = f.BlockCols = f.HandoffLag = 0
    f.autotune(C.DT)
    f.gs_layout()
    f.close()

    # staged host copies (fields above the 2 MB staging threshold), both ways
    f = Fluid(1000.0, 1100, 600, 0.01)
    st = f.state()
    st["U"] = rng.standard_normal(st["U"].shape).astype(np.float32)
    f.load_state(st)
    for k in ("U", "V", "p", "M"):
        f.get(k)
    f.close()

    # a batch of uneven members, eager and graph launches, its tuner and multigrid
    sizes = [(40, 30), (64, 64), (90, 50), (16, 120)]
    b = FluidBatch(C.DENSITY, sizes, C.H)
    for e, m in enumerate(b.members):
        m.reset_to_preset(("jet", "cavity", "karman", "jet")[e], 2.0)
        m.PressureIters = 20
    b.PressureIters = 20
    b.update(C.DT)
    b.set_graph(True)
    b.update(C.DT)
    b.update(C.DT)
    b.set_graph(False)
    b.autotune(C.DT)
    b.UseMultigrid, b.MultigridLevels = True, 2
    b.update(C.DT)
    for m in b.members:
        m.get("U")
    b.close()

    # the rollout adjoint
    f = C.diff_scene(lambda *a: Fluid(*a), 64, 64, 20)
    f.Confinement = 0.01
    z = np.zeros((f.NumX, f.NumY), np.float32)
    r = Rollout(f, 6, checkpoint_every=2)
    r.forward(C.DT, z, z, C.gaussian_blobs(f.NumX, f.NumY))
    r.backward()
    r.close()
    f.UseMultigrid = True  # the rollout through the V-cycle
    r = Rollout(f, 3, checkpoint_every=2)
    r.forward(C.DT, z, z, C.gaussian_blobs(f.NumX, f.NumY))
    r.backward()
    r.close()
    f.close()
    print("asan exercise ok", flush=True)


if __name__ == "__main__":
    main()
