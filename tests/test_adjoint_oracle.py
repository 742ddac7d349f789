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
