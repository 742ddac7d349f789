"""The device adjoint (fl_vjp_*) against the oracle's adjoint: the forward rollout is
bit-identical to the float32 oracle's (same loss up to summation order), and dL/dF agrees
with the oracle's FP64 adjoint (itself checked against central finite differences in
test_adjoint_oracle.py) and with its float32 adjoint along the same trajectory. Finite
differences of the float32 forward are too noisy at the step sizes the confinement's
curvature needs, so they are taken in FP64 on the oracle."""
import numpy as np
import pytest

from fluid_amd import Fluid, Rollout
from fluid_amd import configs as C
from oracle.pyoracle import OracleFluid

pytestmark = pytest.mark.gpu


def jet(make, w, h, iters, bfecc, conf, early=False, visc=0.0, disc=False, lag=0):
    f = C.jet_scene(make, w, h, iters, jet_u=1.0, jet_rows=max(2, h // 4))
    f.UseBFECC, f.Confinement, f.EarlyExit, f.ViscosityDiffusion = bfecc, conf, early, visc
    if lag:  # the device's handoff publication (forward, recompute and mirrored solves)
        f.HandoffLag = lag
    if disc:  # more faces the advection leaves alone (they keep the viscosity's output)
        f.SetCircularObstacle(w // 2, h // 2, 4)
    return f


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


CASES = {
    "sl": dict(bfecc=False, conf=0.0),
    "bfecc": dict(bfecc=True, conf=0.0),
    "bfecc_conf": dict(bfecc=True, conf=0.3),
    "sl_visc_disc": dict(bfecc=False, conf=0.0, visc=0.15, disc=True),
    "bfecc_visc": dict(bfecc=True, conf=0.0, visc=0.15),
    "bfecc_visc_disc": dict(bfecc=True, conf=0.0, visc=0.15, disc=True),
    "bfecc_late_handoff": dict(bfecc=True, conf=0.0, lag=2),
}
# Confinement normalises grad|curl| by (|grad|curl|| + 1e-5) (fluid.go:475-477); where |curl|
# peaks that gradient vanishes and the step's Jacobian reaches ~1e9, so float32 and FP64
# trajectories have different derivatives there. With confinement the device adjoint is
# compared with the float32 oracle's adjoint along the identical trajectory only.
FP64_COMPARABLE = {"sl", "bfecc", "sl_visc_disc", "bfecc_visc", "bfecc_visc_disc", "bfecc_late_handoff"}


@pytest.mark.parametrize("case", list(CASES), ids=list(CASES))
@pytest.mark.parametrize("force", ["zero", "random"])
def test_gradient_matches_oracle(case, force):
    w, h, iters, steps = 40, 32, 16, 6
    kw = CASES[case]
    g = jet(lambda *a: Fluid(*a), w, h, iters, **kw)
    o32 = jet(lambda *a: OracleFluid(*a), w, h, iters, **kw)
    o64 = jet(lambda *a: OracleFluid(*a, prec=64), w, h, iters, **kw)
    rng = np.random.default_rng(7)
    shape = (g.NumX, g.NumY)
    scale = 0.0 if force == "zero" else 0.02
    Fu = (scale * rng.standard_normal(shape)).astype(np.float32)
    Fv = (scale * rng.standard_normal(shape)).astype(np.float32)
    T = C.gaussian_blobs(g.NumX, g.NumY, sigma=4.0)
    r = Rollout(g, steps, checkpoint_every=4)
    loss = r.forward(C.DT, Fu, Fv, T)
    gu, gv = r.backward()
    l32, u32, v32 = o32.rollout(steps, C.DT, Fu, Fv, T)
    l64, u64, v64 = o64.rollout(steps, C.DT, Fu.astype(np.float64), Fv.astype(np.float64), T.astype(np.float64))
    assert abs(loss - l32) <= 1e-9 * abs(l32)  # same float32 trajectory
    G, G32, G64 = np.stack([gu, gv]), np.stack([u32, v32]), np.stack([u64, v64])
    assert rel(G, G32) < 1e-4, rel(G, G32)
    if case not in FP64_COMPARABLE:
        r.close()
        return
    assert rel(G, G64) < 1e-3, rel(G, G64)
    for _ in range(8):
        d = rng.standard_normal(G.shape)
        a, b = float((G * d).sum()), float((G64 * d).sum())
        assert abs(a - b) <= 1e-3 * abs(b) + 1e-6 * np.abs(G64).sum()
    r.close()


def test_checkpoint_interval_and_state():
    w, h, iters, steps = 36, 28, 12, 7
    g = jet(lambda *a: Fluid(*a), w, h, iters, True, 0.3)
    g.update(C.DT)  # start from a flowing state
    start = g.state()
    shape = (g.NumX, g.NumY)
    rng = np.random.default_rng(3)
    Fu = (0.01 * rng.standard_normal(shape)).astype(np.float32)
    Fv = (0.01 * rng.standard_normal(shape)).astype(np.float32)
    T = rng.random(shape).astype(np.float32)
    grads, finals = [], []
    for every in (1, 3, 7):
        r = Rollout(g, steps, checkpoint_every=every)
        loss = r.forward(C.DT, Fu, Fv, T)
        fin = g.state()
        grads.append(np.stack(r.backward()))
        finals.append((loss, fin, g.state()))
        r.close()
        g.load_state(start)
    for loss, fin, after in finals:
        assert loss == finals[0][0]
        for k in fin:  # forward leaves the final state; backward restores it
            np.testing.assert_array_equal(fin[k].view(np.uint32), finals[0][1][k].view(np.uint32))
            np.testing.assert_array_equal(after[k].view(np.uint32), fin[k].view(np.uint32))
    for G in grads[1:]:
        assert rel(G, grads[0]) < 1e-5


def test_early_exit_sweeps_follow_the_forward():
    """A quiescent start makes the early exit keep fewer sweeps in early steps; the
    adjoint must transpose exactly the sweeps that ran (as the float32 oracle does)."""
    w, h, iters, steps = 32, 24, 30, 4
    g = jet(lambda *a: Fluid(*a), w, h, iters, False, 0.0, early=True)
    o32 = jet(lambda *a: OracleFluid(*a), w, h, iters, False, 0.0, early=True)
    shape = (g.NumX, g.NumY)
    z = np.zeros(shape, dtype=np.float32)
    T = C.gaussian_blobs(g.NumX, g.NumY, sigma=3.0)
    r = Rollout(g, steps, checkpoint_every=2)
    loss = r.forward(C.DT, z, z, T)
    G = np.stack(r.backward())
    l32, u32, v32 = o32.rollout(steps, C.DT, z, z, T)
    assert abs(loss - l32) <= 1e-9 * abs(l32)
    assert rel(G, np.stack([u32, v32])) < 1e-4
    r.close()


MG_DENSITY = C.DT / C.H  # cp = density * h / dt = 1


@pytest.mark.parametrize("case", ["sl", "bfecc", "bfecc_disc", "sl_odd"])
def test_gradient_through_multigrid_matches_oracle(case):
    """UseMultigrid: the V-cycle's adjoint (smoothing sweeps carrying p's adjoint, correction,
    prolongation, 40 coarse sweeps, restriction and residual; fluid.go:560-758, 1123-1149
    transposed) against the oracle's, at density dt/h. There cp = 1 and the reference's cycle
    converges; at density 1000 its correction overshoots, the flow saturates the sampler's
    clamps and the solve's adjoint is zero. With the disc, float32 and FP64 trajectories part
    at a kink (their adjoints differ by 1e-2), so that case is compared in float32 only. The
    odd case (39 x 25 with the ring) has odd coarse sizes and odd last rows and columns."""
    w, h, iters, steps = (37, 23, 4, 6) if case == "sl_odd" else (40, 32, 4, 6)
    bfecc, disc = case.startswith("bfecc"), case == "bfecc_disc"

    def make(cls, **kw):
        return lambda _density, *a: cls(MG_DENSITY, *a, **kw)

    fs = [jet(m, w, h, iters, bfecc, 0.0, disc=disc) for m in (make(Fluid), make(OracleFluid),
                                                                 make(OracleFluid, prec=64))]
    for f in fs:
        f.UseMultigrid, f.MultigridLevels = True, 2
    g, o32, o64 = fs
    rng = np.random.default_rng(7)
    shape = (g.NumX, g.NumY)
    Fu = (0.02 * rng.standard_normal(shape)).astype(np.float32)
    Fv = (0.02 * rng.standard_normal(shape)).astype(np.float32)
    T = C.gaussian_blobs(g.NumX, g.NumY, sigma=4.0)
    r = Rollout(g, steps, checkpoint_every=4)
    loss = r.forward(C.DT, Fu, Fv, T)
    gu, gv = r.backward()
    r.close()
    l32, u32, v32 = o32.rollout(steps, C.DT, Fu, Fv, T)
    assert o32.last_sweeps() == 6 * iters  # full cycles: the transposed path has every pass
    assert abs(loss - l32) <= 1e-9 * abs(l32)  # same float32 trajectory
    G, G32 = np.stack([gu, gv]), np.stack([u32, v32])
    assert np.abs(G32).max() > 0
    assert rel(G, G32) < 1e-4, rel(G, G32)
    if disc:
        return
    l64, u64, v64 = o64.rollout(steps, C.DT, Fu.astype(np.float64), Fv.astype(np.float64), T.astype(np.float64))
    G64 = np.stack([u64, v64])
    assert rel(G, G64) < 1e-3, rel(G, G64)


def test_gradient_through_multigrid_exit_matches_oracle():
    """A flow whose divergence stays below 1e-5 (jet speed 1e-6, force 1e-7): each V-cycle
    stops at the exit test after its first pre-smoothing (3 sweeps, fluid.go:576-578), and
    the device transposes only those sweeps."""
    w, h, iters, steps = 40, 32, 4, 4
    fs = []
    for m in (lambda _d, *a: Fluid(MG_DENSITY, *a), lambda _d, *a: OracleFluid(MG_DENSITY, *a)):
        f = C.jet_scene(m, w, h, iters, jet_u=1e-6, jet_rows=8)
        f.UseBFECC, f.Confinement, f.EarlyExit, f.TurbulenceStrength = False, 0.0, False, 0.0
        f.UseMultigrid, f.MultigridLevels = True, 2
        fs.append(f)
    g, o32 = fs
    rng = np.random.default_rng(7)
    shape = (g.NumX, g.NumY)
    Fu = (1e-7 * rng.standard_normal(shape)).astype(np.float32)
    Fv = (1e-7 * rng.standard_normal(shape)).astype(np.float32)
    T = C.gaussian_blobs(g.NumX, g.NumY, sigma=4.0)
    r = Rollout(g, steps, checkpoint_every=2)
    loss = r.forward(C.DT, Fu, Fv, T)
    gu, gv = r.backward()
    r.close()
    l32, u32, v32 = o32.rollout(steps, C.DT, Fu, Fv, T)
    assert o32.last_sweeps() == 3
    assert abs(loss - l32) <= 1e-9 * abs(l32)
    G, G32 = np.stack([gu, gv]), np.stack([u32, v32])
    assert np.abs(G32).max() > 0
    assert rel(G, G32) < 1e-4, rel(G, G32)
