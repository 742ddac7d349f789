"""PreciseAdvection on the device (precise_advection, no reference counterpart): every pass
against the float32 oracle's restatement of it, bit for bit; whole steps, batches and the
adjoint the same way; and the float32 step's error against the FP64 model from the same
state at a size where the reference's absolute-position arithmetic is far off."""
import numpy as np
import pytest

from fluid_amd import Fluid, FluidBatch, Rollout
from fluid_amd import configs as C
from oracle.pyoracle import OracleFluid

from test_gpu_parity import assert_same, random_state, load_both, knobs
from test_gpu_batch import same, scene, set_knobs
from test_precise_oracle import exact_position_state

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("w,h", [(5, 3), (37, 23), (130, 70)])
@pytest.mark.parametrize("name,dt", [("advectVelocity", 0.03), ("advectSmoke", 0.04),
                                     ("advectVelocityBFECC", 0.03), ("advectSmokeBFECC", 0.05)])
@pytest.mark.parametrize("scale", [1.0, 40.0], ids=["inside", "clamped"])
def test_precise_pass_bit_exact(w, h, name, dt, scale):
    g, o = Fluid(3.0, w, h, 0.1), OracleFluid(3.0, w, h, 0.1)
    load_both(g, o, random_state(w + 2, h + 2, seed=w * 7 + h, scale=scale))
    for f in (g, o):
        knobs(f, SmokeAdvection=0.9, PreciseAdvection=True)
        f.run_pass(name, dt)
    assert_same(g, o, name)


@pytest.mark.parametrize("name", ["advectVelocity", "advectSmoke"])
def test_precise_pass_on_exact_positions(name):
    """Back-traces landing exactly on the clamp bounds and beyond (h = 1, dt = 1/4)."""
    st = exact_position_state(26, 19, 1.0, 0.25, seed=7)
    g, o = Fluid(1000.0, 24, 17, 1.0), OracleFluid(1000.0, 24, 17, 1.0)
    load_both(g, o, st)
    for f in (g, o):
        f.PreciseAdvection = True
        f.run_pass(name, 0.25)
    assert_same(g, o, name)


SCENES = {
    "jet_sl": lambda mk: C.jet_scene(mk, 96, 64, 30, jet_rows=8),
    "cavity_bfecc": lambda mk: C.cavity_scene(mk, 128, 96, 60, lid_u=3.0),
    "karman_conf": lambda mk: C.karman_scene(mk, 160, 64, 40, radius=6, centre=(32, 32), band=8),
}


@pytest.mark.parametrize("name", list(SCENES))
def test_precise_steps_bit_exact(name):
    g, o = SCENES[name](lambda *a: Fluid(*a)), SCENES[name](lambda *a: OracleFluid(*a))
    for f in (g, o):
        f.PreciseAdvection = True
    for step in range(6):
        g.update(C.DT)
        o.update(C.DT)
        assert_same(g, o, "%s step %d" % (name, step))


def test_precise_toggles_between_steps():
    """Switching the mode between steps takes effect on the next step, both ways."""
    g, o = (SCENES["cavity_bfecc"](lambda *a: mk(*a)) for mk in (Fluid, OracleFluid))
    for step, precise in enumerate((False, True, True, False, True)):
        for f in (g, o):
            f.PreciseAdvection = precise
            f.update(C.DT)
        assert_same(g, o, "step %d precise %s" % (step, precise))


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
def test_precise_batch_matches_solo(graph):
    sizes = [(64, 48), (33, 77), (96, 96)]
    kn = dict(UseBFECC=True, PressureIters=20, PreciseAdvection=True)
    b = FluidBatch(C.DENSITY, sizes, C.H)
    solo = [Fluid(C.DENSITY, w, h, C.H) for (w, h) in sizes]
    for e, (m, s) in enumerate(zip(b.members, solo)):
        scene(e, m)
        scene(e, s)
        for f in (m, s):
            set_knobs(f, kn)
            f.PreciseAdvection = True
    set_knobs(b, kn)
    b.PreciseAdvection = True
    b.set_graph(graph)
    for step in range(3):
        b.update(C.DT)
        for s in solo:
            s.update(C.DT)
        for e, (m, s) in enumerate(zip(b.members, solo)):
            same(m, s, "step %d member %d" % (step, e))
    # the batch's mode really is precise: turning it off changes the next step
    for s in solo:
        s.PreciseAdvection = False
        s.update(C.DT)
    b.update(C.DT)
    assert any(not np.array_equal(m.get("U"), s.get("U")) for m, s in zip(b.members, solo))
    b.close()


@pytest.mark.parametrize("bfecc", [False, True], ids=["sl", "bfecc"])
def test_precise_gradient_matches_oracle(bfecc):
    w, h, iters, steps = 40, 32, 16, 6

    def jet(mk):
        f = C.jet_scene(mk, w, h, iters, jet_u=1.0, jet_rows=8)
        f.UseBFECC, f.EarlyExit, f.PreciseAdvection = bfecc, False, True
        return f
    g, o32, o64 = jet(lambda *a: Fluid(*a)), jet(lambda *a: OracleFluid(*a)), jet(lambda *a: OracleFluid(*a, prec=64))
    rng = np.random.default_rng(7)
    shape = (g.NumX, g.NumY)
    Fu = (0.02 * rng.standard_normal(shape)).astype(np.float32)
    Fv = (0.02 * rng.standard_normal(shape)).astype(np.float32)
    T = C.gaussian_blobs(g.NumX, g.NumY, sigma=4.0)
    r = Rollout(g, steps, checkpoint_every=4)
    loss = r.forward(C.DT, Fu, Fv, T)
    G = np.stack(r.backward())
    l32, u32, v32 = o32.rollout(steps, C.DT, Fu, Fv, T)
    l64, u64, v64 = o64.rollout(steps, C.DT, Fu.astype(np.float64), Fv.astype(np.float64), T.astype(np.float64))
    assert abs(loss - l32) <= 1e-9 * abs(l32)
    nrm = lambda a, b: float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))  # noqa: E731
    assert nrm(G, np.stack([u32, v32])) < 1e-4
    assert nrm(G, np.stack([u64, v64])) < 1e-3
    r.close()


def test_precise_step_error_against_fp64_cavity_512():
    """One device step from a developed 512^2 cavity state (BFECC, 200 sweeps) against the
    FP64 model's step from the same state: within 1e-5 relative max-abs with PreciseAdvection,
    while the reference arithmetic is several times further off."""
    g = C.cavity_scene(lambda *a: Fluid(*a), 512, 512, 200)
    for _ in range(10):
        g.update(C.DT)
    st = g.state()
    o32 = C.cavity_scene(lambda *a: OracleFluid(*a, threads=16), 512, 512, 200)
    o32.load_state(st)
    o32.update(C.DT)  # sweep count and turbulence gate of the float32 step (fixed before advection)
    o64 = C.cavity_scene(lambda *a: OracleFluid(*a, prec=64, threads=16), 512, 512, 200)
    o64.load_state(st)
    o64.set_forced_sweeps(o32.last_sweeps())
    o64.set_turb_gate(o32.turb_gate())
    o64.update(C.DT)
    errs = {}
    for precise in (False, True):
        g.load_state(st)
        g.PreciseAdvection = precise
        g.update(C.DT)
        errs[precise] = {k: rel(g.get(k), o64.get(k)) for k in ("U", "V", "p", "M")}
    for k in ("U", "V", "p", "M"):
        assert errs[True][k] < 1e-5, (k, errs)
    for k in ("U", "V", "M"):
        assert errs[False][k] > 3 * errs[True][k], (k, errs)
