"""GPU parity: libfluidhip (through the C ABI) against the float32 oracle, bit for bit.

Every test starts the device and the oracle from the same state (uploaded arrays or the
same API call sequence) and compares all eight reference buffers exactly.
"""
import numpy as np
import pytest

from fluid_amd import Fluid
from fluid_amd import configs as C
from known_answers import load_cases, run_case
from oracle.pyoracle import OracleFluid

pytestmark = pytest.mark.gpu

KNOWN = load_cases()


@pytest.mark.parametrize("case", KNOWN, ids=[c["name"] for c in KNOWN])
def test_reference_known_answers_device(case):
    """The reference tests' known answers (tests/golden/reference_known_answers.json) on the
    device library, through the same runner the oracle's CPU test uses."""
    n, fails = run_case(lambda *a: Fluid(*a), case)
    assert n > 0 and not fails, fails

FIELDS = ("U", "V", "newU", "newV", "p", "S", "M", "newM")


def pair(density, w, h, hh):
    return Fluid(density, w, h, hh), OracleFluid(density, w, h, hh)


def assert_same(g, o, where="", fields=FIELDS, nan_equal=False):
    """Bit equality; with nan_equal, NaNs must sit at the same cells but their payloads (which
    x86 and the GPU generate differently) are not compared."""
    for k in fields:
        a, b = g.get(k), o.get(k)
        if nan_equal:
            na, nb = np.isnan(a), np.isnan(b)
            if np.array_equal(na, nb):
                a, b = np.where(na, np.float32(0), a), np.where(nb, np.float32(0), b)
        if not np.array_equal(a.view(np.uint32), b.view(np.uint32)):
            bad = np.argwhere(a.view(np.uint32) != b.view(np.uint32))
            i, j = bad[0]
            raise AssertionError("%s %s differs at %d cells, first (%d,%d): gpu %r oracle %r" %
                                 (where, k, len(bad), i, j, a[i, j], b[i, j]))


def random_state(nx, ny, seed, solid_frac=0.15, walls=True, scale=1.0):
    rng = np.random.default_rng(seed)
    S = (rng.random((nx, ny)) > solid_frac).astype(np.float32)
    if walls:
        S[0, :] = S[:, 0] = S[:, -1] = 0
    st = {"S": S}
    for k in ("U", "V", "newU", "newV"):
        st[k] = (scale * rng.standard_normal((nx, ny))).astype(np.float32)
    st["p"] = rng.standard_normal((nx, ny)).astype(np.float32)
    st["M"] = rng.random((nx, ny)).astype(np.float32)
    st["newM"] = rng.random((nx, ny)).astype(np.float32)
    return st


def load_both(g, o, st):
    g.load_state(st)
    o.load_state(st)


def knobs(f, **kw):
    for k, v in kw.items():
        setattr(f, k, v)


SIZES = [(5, 3), (37, 23), (64, 64), (130, 70)]


@pytest.mark.parametrize("w,h", SIZES)
@pytest.mark.parametrize("name,dt,iters", [
    ("copyBorderUV", 0.0, 0), ("applyViscosity", 0.05, 0), ("applyVorticityConfinement", 0.07, 0),
    ("addTurbulence", 0.1, 0), ("handleBorders", 0.0, 0), ("advectVelocity", 0.03, 0),
    ("advectSmoke", 0.04, 0), ("advectVelocityBFECC", 0.03, 0), ("advectSmokeBFECC", 0.05, 0),
    ("makeIncompressible", 0.1, 1), ("makeIncompressible", 0.1, 13), ("clearP", 0.0, 0)])
def test_pass_bit_exact(w, h, name, dt, iters):
    g, o = pair(3.0, w, h, 0.1)
    st = random_state(w + 2, h + 2, seed=w * 131 + h + iters)
    load_both(g, o, st)
    for f in (g, o):
        knobs(f, ViscosityDiffusion=0.12, Confinement=0.4, TurbulenceStrength=0.5, SmokeAdvection=0.9,
              PressureDamping=0.8, Relaxation=1.7)
        f.run_pass(name, dt, iters)
    assert_same(g, o, name)


@pytest.mark.parametrize("w,h", SIZES)
@pytest.mark.parametrize("iters", [1, 13, 40])
def test_gs_late_handoff_bit_exact(w, h, iters):
    """The temporal-blocked solve with its handoff rows published a block late (HandoffLag 2),
    which the default picks only on large grids, on small and single-tile grids."""
    g, o = pair(3.0, w, h, 0.1)
    load_both(g, o, random_state(w + 2, h + 2, seed=w * 17 + h + iters))
    for f in (g, o):
        knobs(f, PressureDamping=0.8, Relaxation=1.7, EarlyExit=False)
    g.Solver, g.HandoffLag = 2, 2
    for f in (g, o):
        f.run_pass("makeIncompressible", 0.1, iters)
    assert_same(g, o, "late handoff %dx%d, %d sweeps" % (w, h, iters))


@pytest.mark.parametrize("solver", [1, 2], ids=["hyper", "wave"])
@pytest.mark.parametrize("opts", [
    dict(), dict(UseBFECC=True), dict(Confinement=0.3, UseBFECC=True),
    dict(ViscosityDiffusion=0.1, PressureDamping=0.9, Relaxation=1.5),
    dict(TurbulenceStrength=0.0, SmokeAdvection=1.3, EarlyExit=False)])
def test_simulate_bit_exact(solver, opts):
    g, o = pair(1000.0, 45, 33, 0.01)
    st = random_state(47, 35, seed=7, scale=0.6)
    load_both(g, o, st)
    for f in (g, o):
        knobs(f, PressureIters=17, **opts)
    g.Solver = solver
    for step in range(4):
        g.Simulate(0.1)
        o.Simulate(0.1)
        assert_same(g, o, "step %d" % step)
    assert g.stats()["sweeps_run"] == o.last_sweeps()


@pytest.mark.parametrize("shape", [(4, 8), (8, 8), (16, 4)], ids=["tk4", "tk8", "tk16_blk4"])
def test_early_exit_after_first_sweep_repeated(shape):
    """Steps whose solve ends after sweep 0, one after another: once a solve has ended there,
    the next main run also keeps its first sweep's result and takes it instead of a re-run.
    Every step is bit-exact with the oracle and reports one kept sweep below the tolerance."""
    g, o = pair(1000.0, 70, 50, 0.01)
    st = {k: np.zeros((72, 52), np.float32) for k in FIELDS}
    st["S"] = np.ones((72, 52), np.float32)
    st["S"][0, :] = st["S"][:, 0] = st["S"][:, -1] = 0
    load_both(g, o, st)
    for f in (g, o):
        knobs(f, PressureIters=20)
    g.BlockSweeps, g.BlockCols = shape
    for step in range(5):
        U = np.zeros((72, 52), np.float32)
        U[30 + step, 20] = 2e-8 * (step + 1)  # divergence below the tolerance after one sweep
        for f in (g, o):
            f.set("U", U)
            f.Simulate(0.1)
        assert_same(g, o, "step %d" % step)
        s = g.stats()  # also waits for the step, so the next solve sees how this one ended
        assert s["sweeps_run"] == o.last_sweeps() == 1 and s["reran"], (step, s)
        assert 0.0 <= s["max_div_last"] < 1e-5, s


@pytest.mark.parametrize("solver", [1, 2], ids=["hyper", "wave"])
@pytest.mark.parametrize("lag", [0, 2], ids=["auto_lag", "late_handoff"])
def test_early_exit_rollback(solver, lag):
    """A quiescent field converges in sweep 0: the device must keep exactly one sweep."""
    g, o = pair(1000.0, 40, 30, 0.01)
    st = {k: np.zeros((42, 32), np.float32) for k in FIELDS}
    st["S"] = np.ones((42, 32), np.float32)
    st["S"][0, :] = 0
    load_both(g, o, st)
    for f in (g, o):
        knobs(f, PressureIters=25)
    g.Solver, g.HandoffLag = solver, lag
    # small divergence that the first sweeps remove below 1e-5 after a few sweeps
    U = np.zeros((42, 32), np.float32)
    U[20, 15] = 3e-5
    g.set("U", U)
    o.set("U", U)
    g.Simulate(0.1)
    o.Simulate(0.1)
    assert_same(g, o)
    s = g.stats()
    assert s["sweeps_run"] == o.last_sweeps() < 25 and s["reran"]


def test_jet_256_200_steps():
    """The 256^2 jet (40 sweeps, SL, dt 0.1) for 200 steps, compared every step."""
    g = C.jet_scene(lambda *a: Fluid(*a))
    o = C.jet_scene(lambda *a: OracleFluid(*a))
    assert_same(g, o, "setup")
    for step in range(200):
        g.update(C.DT)
        o.update(C.DT)
        fields = FIELDS if step % 20 == 19 else ("U", "V", "p", "M")
        assert_same(g, o, "step %d" % step, fields)
        assert g.stats()["sweeps_run"] == 40


@pytest.mark.parametrize("preset", ["jet", "cavity", "karman"])
def test_presets_and_update(preset):
    g, o = pair(1000.0, 70, 50, 0.01)
    for f in (g, o):
        f.reset_to_preset(preset, 3.0)
        if preset == "jet":
            f.set_jet(True, 1, 20, 30, 4.0, 1.0)
        f.add_sink(30, 25, 0.0, -1.0)
        knobs(f, PressureIters=12, UseBFECC=(preset == "karman"))
    assert_same(g, o, "preset")
    for step in range(6):
        g.update(1.0 / 120.0)
        o.update(1.0 / 120.0)
        assert_same(g, o, "%s step %d" % (preset, step))


def test_scaled_configs_match():
    """Karman and cavity builders at reduced size, BFECC and confinement on."""
    for build in (lambda mk: C.karman_scene(mk, 200, 100, 30, radius=12, centre=(40, 50)),
                  lambda mk: C.cavity_scene(mk, 160, 160, 30)):
        g = build(lambda *a: Fluid(*a))
        o = build(lambda *a: OracleFluid(*a))
        assert_same(g, o, "setup")
        for step in range(5):
            g.update(C.DT)
            o.update(C.DT)
            assert_same(g, o, "step %d" % step)


def test_edit_api_sequence():
    g, o = pair(1.0, 30, 20, 1.0)
    st = random_state(32, 22, seed=11, walls=False, solid_frac=0.0)
    load_both(g, o, st)
    for f in (g, o):
        f.SetSolid(3, 4, True)
        f.SetVelocity(4, 4, 2.5, -1.5)
        f.AddSmoke(4, 4, 0.25)
        f.AddSmoke(4, 4, 0.125)
        f.ApplyForce(4, 4, 1.0, 1.0)
        f.ApplyForce(3, 4, 9.0, 9.0)   # solid: ignored
        f.ApplyForceRadius(10, 10, 3.0, -2.0, 4)
        f.SetCircularObstacle(20, 12, 3)
        f.SetSolid(20, 12, False)
        f.ApplyForceRadius(20, 12, 1.0, 1.0, 0)
        f.draw_line(1, 1, 25, 17, True)
        f.smoke_brush(8, 15, 3, 2.0)
        f.SetVelocity(0, 0, 5.0, 5.0)
    assert_same(g, o, "edits")
    assert g.IsSolid(20, 13) == o.IsSolid(20, 13)
    for f in (g, o):
        f.Reset()
    assert_same(g, o, "reset")
    with pytest.raises(IndexError, match="invalid x-index: 32"):
        g.SetSolid(32, 0, True)
    with pytest.raises(IndexError, match="invalid y-index: 22"):
        g.IsSolid(0, 22)


def test_views_and_reductions():
    g, o = pair(1000.0, 50, 40, 0.01)
    st = random_state(52, 42, seed=13, scale=2.0)
    load_both(g, o, st)
    for name in ("Vorticity", "VelocityMagnitude", "Pressure", "Smoke"):
        a, b = getattr(g, name)(), getattr(o, name)()
        assert np.array_equal(a.values, b.values), name
        assert a.MinValue == np.float32(b.MinValue) and a.MaxValue == np.float32(b.MaxValue), name
    assert g.MaxDivergence() == np.float32(o.MaxDivergence())
    for base in (0.1, 1.0 / 120.0):
        assert g.GetAdaptiveTimeStep(base) == np.float32(o.GetAdaptiveTimeStep(base))
    rng = np.random.default_rng(5)
    xs = (rng.random(500) * 0.6 - 0.05).astype(np.float32)
    ys = (rng.random(500) * 0.5 - 0.05).astype(np.float32)
    us, vs = g.SampleVelocityBatch(xs, ys)
    for k in range(0, 500, 7):
        u, v = o.SampleVelocity(float(xs[k]), float(ys[k]))
        assert us[k] == np.float32(u) and vs[k] == np.float32(v)
    assert g.SampleVelocity(0.2, 0.15) == tuple(np.float32(x) for x in o.SampleVelocity(0.2, 0.15))


def test_fp64_per_step_error():
    """Per-step error of the device FP32 step against the FP64 model from the same state
    (config-1 geometry at 128^2 after 30 steps of flow), relative max-abs <= 1e-5."""
    g = C.jet_scene(lambda *a: Fluid(*a), 128, 128, 40)
    for _ in range(30):
        g.update(C.DT)
    st = g.state()
    d = OracleFluid(1000.0, 128, 128, 0.01, prec=64)
    d.load_state(st)
    knobs(d, PressureIters=40)
    d.set_jet(True, 1, 56, 72, 1.0, 1.0)
    o32 = OracleFluid(1000.0, 128, 128, 0.01)
    o32.load_state(st)
    knobs(o32, PressureIters=40)
    o32.set_jet(True, 1, 56, 72, 1.0, 1.0)
    o32.update(C.DT)
    g.update(C.DT)
    d.set_forced_sweeps(o32.last_sweeps())
    d.set_turb_gate(o32.turb_gate())
    d.update(C.DT)
    for k in ("U", "V", "p", "M"):
        a, b = g.get(k).astype(np.float64), d.get(k)
        err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)
        assert err <= 1e-5, (k, err)


def test_fill_cells_matches_direct_writes():
    """fl_fill_cells against the same writes on the oracle's fields: after a queued edit, on
    a column, a whole row of the last column and rows of V; an empty range writes nothing; S
    and ranges outside the grid are refused."""
    from fluid_amd._lib import FluidError
    g, o = pair(1000.0, 37, 23, 0.01)
    load_both(g, o, random_state(39, 25, seed=11))
    for f in (g, o):
        f.SetVelocity(1, 7, 1.0, 2.0)
        f.fill_cells("U", 1, 2, 5, 15, 4.0)
        f.fill_cells("M", 0, 39, 24, 25, 0.5)
        f.fill_cells("V", 10, 12, 0, 25, -1.0)
        f.fill_cells("p", 3, 3, 0, 5, 9.0)
    assert_same(g, o, "fills")
    g.Simulate(0.1)
    o.Simulate(0.1)
    assert_same(g, o, "step after fills")
    for bad in (("S", 0, 1, 0, 1), ("U", 0, 40, 0, 1), ("U", -1, 1, 0, 1), ("M", 0, 1, 0, 26)):
        with pytest.raises(FluidError):
            g.fill_cells(*bad, 0.0)


@pytest.mark.parametrize("w,h", [(1100, 700), (2500, 1700)])
def test_large_transfers_round_trip(w, h):
    """Uploads and downloads of 2 MB and more go through pinned staging in 8 MB chunks: one
    partial chunk (1100x700, 3.1 MB per field) and three (2500x1700, 17 MB) round-trip bit for
    bit, and a download shows what device-side writes left."""
    f = Fluid(1000.0, w, h, 0.01)
    a = np.random.default_rng(3).standard_normal((f.NumX, f.NumY)).astype(np.float32)
    f.set("U", a)
    assert np.array_equal(f.get("U").view(np.uint32), a.view(np.uint32))
    f.fill_cells("M", 0, f.NumX, 0, f.NumY, 0.25)
    f.fill_cells("M", 7, 9, 3, f.NumY - 1, -1.5)
    b = np.full((f.NumX, f.NumY), 0.25, np.float32)
    b[7:9, 3:f.NumY - 1] = -1.5
    assert np.array_equal(f.get("M"), b)
    f.close()


# the coarse solve runs all 40 sweeps in one launch up to a 64 x 64 coarse interior (fine
# interiors up to 129 x 129), and by tile diagonals beyond (131 x 62 and 62 x 131 are one coarse
# row or column over)
MG_SIZES = [(5, 3), (6, 4), (37, 23), (64, 64), (101, 60), (129, 129), (131, 62), (62, 131)]


@pytest.mark.parametrize("w,h", MG_SIZES)
@pytest.mark.parametrize("bfecc", [False, True], ids=["sl", "bfecc"])
@pytest.mark.parametrize("scale", [0.6, 0.02], ids=["strong", "mild"])
def test_multigrid_bit_exact(w, h, bfecc, scale):
    """UseMultigrid (fluid.go:148-151, 559-758): pre/post smoothing, residual, full-weighting
    restriction, 40 coarse sweeps, bilinear prolongation and the serial correction. The
    reference's V-cycle diverges on the strong random states (NaN after a few steps, at the
    same cells on both sides); the mild ones stay finite, but for 0.5% of the U faces of the
    129 x 129 grid in its third step (on both sides alike)."""
    g, o = pair(1000.0, w, h, 0.01)
    st = random_state(w + 2, h + 2, seed=w * 7 + h, scale=scale)
    load_both(g, o, st)
    for f in (g, o):
        knobs(f, UseMultigrid=True, MultigridLevels=2, PressureIters=4, UseBFECC=bfecc, Confinement=0.2)
    for step in range(3):
        g.Simulate(0.1)
        o.Simulate(0.1)
        assert_same(g, o, "step %d" % step, nan_equal=True)
    if scale < 0.1:
        assert np.isfinite(g.get("U")).mean() > 0.99 and np.isfinite(g.get("p")).all()
    assert g.stats()["sweeps_run"] == o.last_sweeps()


@pytest.mark.parametrize("w,h", [(260, 131), (300, 190)])
@pytest.mark.parametrize("scene", ["cavity", "jet"])
def test_multigrid_scenes_bit_exact(w, h, scene):
    """The V-cycle on coarse grids of several tiles in both directions (260x131 leaves a 1-row
    and a 1-column edge tile), from the presets: the random states diverge at these sizes."""
    mk = C.cavity_scene if scene == "cavity" else C.jet_scene
    g, o = mk(lambda *a: Fluid(*a), w, h, 4), mk(lambda *a: OracleFluid(*a), w, h, 4)
    for f in (g, o):
        f.UseMultigrid, f.MultigridLevels = True, 2
    for step in range(6):
        g.update(C.DT)
        o.update(C.DT)
        assert_same(g, o, "%s step %d" % (scene, step))
    assert np.isfinite(g.get("U")).all() and np.isfinite(g.get("p")).all()
    assert g.stats()["sweeps_run"] == o.last_sweeps()


def test_multigrid_pass_and_converged_break():
    # the stand-alone makeIncompressible pass with the V-cycle
    g, o = pair(3.0, 50, 41, 0.1)
    load_both(g, o, random_state(52, 43, seed=5))
    for f in (g, o):
        knobs(f, UseMultigrid=True, MultigridLevels=3, PressureDamping=0.7)
        f.run_pass("makeIncompressible", 0.1, 5)
    assert_same(g, o, "pass")
    # a quiescent field meets the tolerance after the first pre-smoothing: 3 sweeps, no V-cycle
    g, o = pair(1000.0, 30, 20, 0.01)
    st = {k: np.zeros((32, 22), np.float32) for k in FIELDS}
    st["S"] = np.ones((32, 22), np.float32)
    st["S"][0, :] = 0
    load_both(g, o, st)
    for f in (g, o):
        knobs(f, UseMultigrid=True, PressureIters=6)
        f.Simulate(0.1)
    assert_same(g, o, "quiescent")
    assert g.stats()["sweeps_run"] == o.last_sweeps() == 3


def test_field_ranges_follow_the_scan():
    """Pressure()/Smoke() ranges (smoke.go:6-15) on the device: the first element in index
    order wins among equal values (+0 vs -0), NaN is never selected, +-MaxFloat32 when
    nothing beats the initial bounds."""
    g, o = pair(1000.0, 40, 30, 0.01)
    rng = np.random.default_rng(3)
    M = rng.standard_normal((42, 32)).astype(np.float32)
    p = np.zeros((42, 32), np.float32)
    p[5, 5], p[7, 3], p[9, 1] = -0.0, 0.0, np.nan   # all-zero field: the first zero is -0
    M[3, 4] = np.nan
    for f in (g, o):
        f.set("M", M)
        f.set("p", p)
    for fn in ("Smoke", "Pressure"):
        a, b = getattr(g, fn)(), getattr(o, fn)()
        gm = np.array([a.MinValue, a.MaxValue], dtype=np.float32)
        om = np.array([b.MinValue, b.MaxValue], dtype=np.float32)
        assert gm.view(np.uint32).tolist() == om.view(np.uint32).tolist(), (fn, gm, om)
    # nothing beats the initial bounds: an all-NaN field
    for f in (g, o):
        f.set("p", np.full((42, 32), np.nan, np.float32))
    a, b = g.Pressure(), o.Pressure()
    assert (a.MinValue, a.MaxValue) == (b.MinValue, b.MaxValue) == (float(np.finfo(np.float32).max),
                                                                     -float(np.finfo(np.float32).max))


@pytest.mark.parametrize("view", [0, 1, 2, 3], ids=["smoke", "pressure", "speed", "vorticity"])
def test_render_matches_oracle(view):
    g = C.jet_scene(lambda *a: Fluid(*a), 90, 70, 20, jet_rows=10)
    o = C.jet_scene(lambda *a: OracleFluid(*a), 90, 70, 20, jet_rows=10)
    for f in (g, o):
        f.SetCircularObstacle(40, 35, 6)
        f.add_source(10, 60, 0.5, 0.0)
        f.add_sink(80, 10, 0.2, 0.1)
        f.add_source(300, 5, 0.0, 0.0)   # outside the grid: its marker index wraps as in the reference
        f.UseBFECC, f.Confinement = True, 0.2
    for _ in range(6):
        g.update(C.DT)
        o.update(C.DT)
    assert_same(g, o, "state")
    a, b = g.render(view), o.render(view)
    assert a.shape == b.shape == (g.NumY, g.NumX, 4)
    bad = np.argwhere((a != b).any(axis=2))
    assert len(bad) == 0, (len(bad), bad[:3], a[tuple(bad[0])], b[tuple(bad[0])])


def test_particles_match_oracle():
    g = C.jet_scene(lambda *a: Fluid(*a), 80, 64, 20, jet_rows=12)
    o = C.jet_scene(lambda *a: OracleFluid(*a), 80, 64, 20, jet_rows=12)
    for f in (g, o):
        f.SetCircularObstacle(30, 32, 8)
    for _ in range(5):
        g.update(C.DT)
        o.update(C.DT)
    rng = np.random.default_rng(9)
    n = 5000
    h = np.float32(C.H)
    x = (rng.random(n) * 84 - 1).astype(np.float32) * h   # some start outside the grid
    y = (rng.random(n) * 68 - 1).astype(np.float32) * h
    age = (rng.random(n) * 1.4).astype(np.float32)
    mx = np.full(n, 1.5, np.float32)
    gx, gy, ga, gm = x.copy(), y.copy(), age.copy(), mx.copy()
    ox, oy, oa, om = x.copy(), y.copy(), age.copy(), mx.copy()
    for frame in range(4):
        kg = g.advect_particles(np.float32(C.DT), gx, gy, ga, gm)
        ko = o.advect_particles(np.float32(C.DT), ox, oy, oa, om)
        assert np.array_equal(kg, ko), frame
        for a, b in ((gx, ox), (gy, oy), (ga, oa)):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), frame
        keep = kg.astype(bool)
        assert 0 < keep.sum() < len(keep)
        gx, gy, ga, gm = [np.ascontiguousarray(a[keep]) for a in (gx, gy, ga, gm)]
        ox, oy, oa, om = [np.ascontiguousarray(a[keep]) for a in (ox, oy, oa, om)]
        g.update(C.DT)
        o.update(C.DT)


def test_context_error_text_survives_thread_switches():
    """Two threads fail calls on two contexts at once; each context keeps the reference's
    exact panic text of its own last failure (walls.go:6-11), readable from any thread, while
    the per-thread fl_last_error of the reading thread is unrelated."""
    import threading
    a, b = Fluid(1.0, 8, 8, 1.0), Fluid(1.0, 8, 8, 1.0)
    go = threading.Barrier(2)

    def fail(f, i, j, reps):
        go.wait()
        for _ in range(reps):
            assert f._L.fl_set_solid(f._ctx, i, j, 1) == (2 if i > 9 else 3)
    ta = threading.Thread(target=fail, args=(a, 17, 1, 500))
    tb = threading.Thread(target=fail, args=(b, 1, 23, 500))
    ta.start()
    tb.start()
    ta.join()
    tb.join()
    assert a.last_error() == "invalid x-index: 17"
    assert b.last_error() == "invalid y-index: 23"
    with pytest.raises(IndexError, match="^invalid x-index: -1$"):
        a.SetVelocity(-1, 2, 0.0, 0.0)
    assert b.last_error() == "invalid y-index: 23"


def test_solid_mask_upload_rejects_other_values():
    """S must hold exactly +0.0 or 1.0 (every writer of the reference produces those), so the
    device's S != 0 tests agree with the reference's S > 0 ones (fluid.go:536): negative,
    -0.0, NaN and fractional values are refused and the mask is left as it was."""
    g = Fluid(1.0, 6, 5, 1.0)
    g.set("S", np.ones((8, 7), np.float32))
    for bad in (-1.0, -0.0, np.nan, 0.5):
        S = np.ones((8, 7), np.float32)
        S[3, 3] = bad
        with pytest.raises(Exception):
            g.set("S", S)
        assert (g.get("S") == 1.0).all()


def test_staged_frame_outputs_match_oracle():
    """The per-frame outputs above the 2 MB staging threshold, whose host copies go through
    the pinned chunked buffers: the image of an 800x700 grid (2.2 MB), 2^19 tracer particles
    (2 MB per array) and 2^19 SampleVelocity points (4 MB of positions), against the oracle."""
    g = C.jet_scene(lambda *a: Fluid(*a), 800, 700, 8, jet_rows=40)
    o = C.jet_scene(lambda *a: OracleFluid(*a, threads=8), 800, 700, 8, jet_rows=40)
    for f in (g, o):
        f.SetCircularObstacle(200, 350, 30)
    for _ in range(3):
        g.update(C.DT)
        o.update(C.DT)
    for view in (0, 3):
        a, b = g.render(view), o.render(view)
        assert np.array_equal(a, b), view
    n = 1 << 19
    rng = np.random.default_rng(12)
    h = np.float32(C.H)
    x = (rng.random(n) * 803 - 1).astype(np.float32) * h
    y = (rng.random(n) * 703 - 1).astype(np.float32) * h
    us, vs = g.SampleVelocityBatch(x, y)
    for k in rng.integers(0, n, 400):
        ou, ov = o.SampleVelocity(float(x[k]), float(y[k]))
        assert us[k] == np.float32(ou) and vs[k] == np.float32(ov), k
    age = (rng.random(n) * 1.4).astype(np.float32)
    mx = np.full(n, 1.5, np.float32)
    gx, gy, ga = x.copy(), y.copy(), age.copy()
    ox, oy, oa = x.copy(), y.copy(), age.copy()
    kg = g.advect_particles(np.float32(C.DT), gx, gy, ga, mx)
    ko = o.advect_particles(np.float32(C.DT), ox, oy, oa, mx.copy())
    assert np.array_equal(kg, ko)
    for a, b in ((gx, ox), (gy, oy), (ga, oa)):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_batch_member_staged_round_trip():
    """A batch member's field of 16 MB and more goes through the batch threshold's staged copy
    (2048x2048 member, 16.8 MB per field), bit for bit both ways, beside a small member."""
    from fluid_amd import FluidBatch
    b = FluidBatch(C.DENSITY, [(2048, 2048), (30, 20)], C.H)
    big = b.members[0]
    a = np.random.default_rng(4).standard_normal((big.NumX, big.NumY)).astype(np.float32)
    big.set("M", a)
    assert np.array_equal(big.get("M").view(np.uint32), a.view(np.uint32))
    small = b.members[1]
    s = np.random.default_rng(5).standard_normal((small.NumX, small.NumY)).astype(np.float32)
    small.set("U", s)
    assert np.array_equal(small.get("U").view(np.uint32), s.view(np.uint32))
    assert np.array_equal(big.get("M").view(np.uint32), a.view(np.uint32))
    b.close()
