"""The oracle's adjoint of the differentiable rollout (oracle/fluid_adjoint.inc) against
central finite differences of its own FP64 forward pass, on random directions. A dropped
term, a wrong sign or a transposed index in any pass's adjoint changes the directional
derivative by far more than the tolerances here. Confinement's normalised gradient of |curl|
is sharply curved (eps = 1e-5 in its denominator), so those cases use a smaller FD step."""
import numpy as np
import pytest

from fluid_amd import configs as C
from oracle.pyoracle import OracleFluid


def scene(w=20, h=16, iters=6, bfecc=True, conf=0.0, turb=0.02, walls=(True, True, True, True), disc=False):
    o = OracleFluid(C.DENSITY, w, h, C.H, prec=64)
    C.jet_scene(lambda *a: o, w, h, iters, jet_u=1.0, jet_rows=4, walls=walls)
    if disc:
        o.SetCircularObstacle(w // 2, h // 2, 2)
    o.UseBFECC, o.Confinement, o.TurbulenceStrength, o.EarlyExit = bfecc, conf, turb, False
    return o


def fd_errors(o, steps, eps, ndir=4, fscale=0.05, seed=0):
    rng = np.random.default_rng(seed)
    shape = (o.NumX, o.NumY)
    Fu, Fv = fscale * rng.standard_normal(shape), fscale * rng.standard_normal(shape)
    T = rng.random(shape)
    loss, gu, gv = o.rollout(steps, C.DT, Fu, Fv, T)
    assert np.isfinite(loss) and np.abs(gu).max() > 0
    errs = []
    for _ in range(ndir):
        du, dv = rng.standard_normal(shape), rng.standard_normal(shape)
        lp = o.rollout(steps, C.DT, Fu + eps * du, Fv + eps * dv, T, grad=False)
        lm = o.rollout(steps, C.DT, Fu - eps * du, Fv - eps * dv, T, grad=False)
        fd = (lp - lm) / (2 * eps)
        ad = (gu * du).sum() + (gv * dv).sum()
        errs.append(abs(fd - ad) / abs(fd))
    return errs


@pytest.mark.parametrize("bfecc", [False, True], ids=["sl", "bfecc"])
@pytest.mark.parametrize("walls,disc", [((True, True, True, True), False), ((True, False, True, False), True)],
                         ids=["box", "open_disc"])
def test_adjoint_matches_finite_differences(bfecc, walls, disc):
    o = scene(bfecc=bfecc, walls=walls, disc=disc)
    assert max(fd_errors(o, 3, 1e-6)) < 1e-5


@pytest.mark.parametrize("bfecc", [False, True], ids=["sl", "bfecc"])
def test_adjoint_with_confinement(bfecc):
    o = scene(bfecc=bfecc, conf=0.3)
    assert max(fd_errors(o, 2, 1e-8)) < 1e-4


def test_adjoint_of_the_pressure_sweeps_alone():
    # no turbulence, no advection of velocity into smoke gradients beyond one step: the
    # sweeps' transpose carries the whole velocity adjoint
    o = scene(bfecc=False, turb=0.0, iters=12)
    assert max(fd_errors(o, 1, 1e-6, ndir=6)) < 1e-6


def test_rollout_leaves_state_and_zero_force_gradient_is_finite():
    o = scene(bfecc=True, conf=0.3)
    st = o.state()
    shape = (o.NumX, o.NumY)
    loss, gu, gv = o.rollout(4, C.DT, np.zeros(shape), np.zeros(shape), np.random.default_rng(1).random(shape))
    for k, v in st.items():
        np.testing.assert_array_equal(o.get(k), v)
    assert np.isfinite(gu).all() and np.isfinite(gv).all() and np.abs(gu).max() > 0
    # the force only acts on fluid interior cells
    S = o.get("S")
    assert not gu[S == 0].any() and not gu[0, :].any() and not gu[:, -1].any()


@pytest.mark.parametrize("bfecc", [False, True], ids=["sl", "bfecc"])
@pytest.mark.parametrize("disc", [False, True], ids=["box", "disc"])
def test_adjoint_with_viscosity(bfecc, disc):
    """Viscosity's Laplacian step transposed, the faces the advection does not advect (they copy
    the viscosity's newU/newV) and the smoke diffusion inside advectSmoke. The step is 1e-7:
    with the diffused smoke, one direction at 1e-6 straddles a kink (the smoke clamp at zero or
    the turbulence's speed gate) and the central difference is off by 2.6e-2 there."""
    o = scene(bfecc=bfecc, disc=disc, walls=(True, False, True, False) if disc else (True, True, True, True))
    o.ViscosityDiffusion = 0.15
    assert max(fd_errors(o, 3, 1e-7)) < 1e-5
    assert max(fd_errors(o, 3, 1e-7, seed=3)) < 1e-5


def mg_scene(iters=4, disc=False, bfecc=False, jet_u=1.0, turb=0.02, prec=64):
    """The multigrid V-cycle (fluid.go:560-599) at density dt/h, so cp = density * h / dt = 1.
    The cycle subtracts lap(p) from the divergence while p accumulates cp * q (fluid.go:601-632,
    1123-1149); at the scenes' density 1000 (cp = 1000) that correction overshoots and the
    divergence grows about tenfold per cycle, the flow saturates the sampler's clamps and the
    loss stops depending on F through the solve (the transposed passes then see no adjoint)."""
    o = OracleFluid(C.DT / C.H, 20, 16, C.H, prec=prec)
    C.jet_scene(lambda *a: o, 20, 16, iters, jet_u=jet_u, jet_rows=4)
    if disc:
        o.SetCircularObstacle(10, 8, 2)
    o.UseBFECC, o.Confinement, o.TurbulenceStrength, o.EarlyExit = bfecc, 0.0, turb, False
    o.UseMultigrid, o.MultigridLevels = True, 2
    return o


@pytest.mark.parametrize("bfecc", [False, True], ids=["sl", "bfecc"])
@pytest.mark.parametrize("disc", [False, True], ids=["box", "disc"])
@pytest.mark.parametrize("steps", [1, 3])
def test_adjoint_through_multigrid(bfecc, disc, steps):
    """The V-cycle's smoothing sweeps with p's adjoint, correction, prolongation, 40 coarse
    sweeps, restriction and residual, each transposed (fluid_adjoint.inc mg_adj)."""
    o = mg_scene(disc=disc, bfecc=bfecc)
    assert max(fd_errors(o, steps, 1e-6)) < 1e-5
    assert o.last_sweeps() == 24  # four full cycles, no exit


def test_adjoint_through_multigrid_exit():
    """A flow with divergence below 1e-5: the exit test after the first pre-smoothing stops the
    V-cycle (3 sweeps), and only those sweeps are transposed."""
    o = mg_scene(iters=3, jet_u=1e-6, turb=0.0)
    assert max(fd_errors(o, 1, 1e-8, fscale=1e-7)) < 1e-5
    assert o.last_sweeps() == 3


@pytest.mark.parametrize("plant", [27, 28, 29])
def test_multigrid_adjoint_pin_catches_planted_mistakes(plant):
    """27: the residual's p term with the opposite sign; 28: the coarse sweep's own (1 - relax)
    factor dropped; 29: the correction's cp * lambda_p term dropped."""
    from oracle import pyoracle
    with pyoracle.planted(plant):
        o = mg_scene(disc=True)
        assert max(fd_errors(o, 2, 1e-6)) > 1e-3
