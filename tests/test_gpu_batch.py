"""Batched ensembles: a FluidBatch steps members of uneven sizes with one launch per pass
and one persistent pressure launch, and must leave every member bit-identical to the same
Fluid stepped on its own (which the parity tests pin to the oracle). The reference's
ensemble is a loop of Simulate calls (fluid.go:79)."""
import numpy as np
import pytest

from fluid_amd import Fluid, FluidBatch
from fluid_amd import configs as C
from oracle.pyoracle import OracleFluid

pytestmark = pytest.mark.gpu

FIELDS = ("U", "V", "newU", "newV", "p", "S", "M", "newM")

# uneven members: widths/heights straddling the 64-lane tile edge, a thin one, a tiny one
SIZES = [(48, 40), (130, 70), (17, 95), (64, 64), (200, 33), (1, 1), (66, 129)]


def same(a, b, where, fields=FIELDS):
    for k in fields:
        x, y = a.get(k), b.get(k)
        bad = x.view(np.uint32) != y.view(np.uint32)
        if bad.any():
            i, j = np.argwhere(bad)[0]
            raise AssertionError("%s: %s differs at %d cells, first (%d,%d) %r vs %r" %
                                 (where, k, int(bad.sum()), i, j, x[i, j], y[i, j]))


def scene(e, f):
    """Member e's scene: jet, cavity, obstacle with inflow, or an idle box (zero
    velocity, so the early exit re-runs a single sweep)."""
    w, h = f.NumX - 2, f.NumY - 2
    kind = e % 4
    mk = lambda *a: f  # noqa: E731
    if kind == 0:
        C.jet_scene(mk, w, h, jet_rows=max(2, h // 8))
    elif kind == 1:
        C.cavity_scene(mk, w, h)
    elif kind == 2:
        C.karman_scene(mk, w, h, radius=max(1, min(w, h) // 6), centre=(max(1, w // 4), max(1, h // 2)),
                       band=max(1, min(4, w)))
    else:
        C.open_all_cells(f)
        f.scene_walls(True, True, True, True)
        f.apply_wall_settings()


BASE_KNOBS = dict(Confinement=0.0, ViscosityDiffusion=0.0, PressureDamping=1.0, TurbulenceStrength=0.02,
                  SmokeAdvection=1.0, UseBFECC=False, PressureIters=8, EarlyExit=True, BlockSweeps=0,
                  HandoffLag=0)


def set_knobs(dst, knobs):
    """Every knob, so that what a scene set on a member does not linger."""
    for k, v in dict(BASE_KNOBS, **knobs).items():
        setattr(dst, k, v)


KNOB_SETS = {
    "bfecc_conf": dict(UseBFECC=True, Confinement=0.3, PressureIters=40, EarlyExit=True),
    "sl_visc": dict(UseBFECC=False, ViscosityDiffusion=0.002, PressureIters=24, EarlyExit=False),
    "tk8_damped": dict(UseBFECC=True, PressureDamping=0.8, PressureIters=20, EarlyExit=True, BlockSweeps=8),
    "late_handoff": dict(UseBFECC=True, PressureIters=40, EarlyExit=True, HandoffLag=2),
}


def make_pair(sizes, knobs):
    b = FluidBatch(C.DENSITY, sizes, C.H)
    solo = [Fluid(C.DENSITY, w, h, C.H) for (w, h) in sizes]
    for e, (m, s) in enumerate(zip(b.members, solo)):
        scene(e, m)
        scene(e, s)
        set_knobs(s, knobs)
        set_knobs(m, knobs)
    set_knobs(b, knobs)
    return b, solo


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
@pytest.mark.parametrize("knobs", list(KNOB_SETS), ids=list(KNOB_SETS))
def test_batch_matches_solo(knobs, graph):
    b, solo = make_pair(SIZES, KNOB_SETS[knobs])
    b.set_graph(graph)
    for step in range(4):
        b.update(C.DT)
        for s in solo:
            s.update(C.DT)
        for e, (m, s) in enumerate(zip(b.members, solo)):
            same(m, s, "%s step %d member %d %s" % (knobs, step, e, SIZES[e]))
    # per-member solve statistics follow each member's own early exit
    for m, s in zip(b.members, solo):
        assert m.stats()["sweeps_run"] == s.stats()["sweeps_run"]
        np.testing.assert_array_equal(m.sweep_maxdiv().view(np.uint32), s.sweep_maxdiv().view(np.uint32))
    b.close()


SHAPES = [(tk, blk) for tk in (2, 4, 8, 16) for blk in (4, 8, 16)]


@pytest.mark.parametrize("shape", SHAPES, ids=["tk%d_blk%d" % s for s in SHAPES])
def test_batch_every_shape_matches_solo(shape):
    """The merged persistent launch with each tile shape the batch autotuner can pick (with a
    deferred handoff publication) leaves every member bit-identical to stepping it alone
    with the default shape."""
    knobs = dict(KNOB_SETS["bfecc_conf"], BlockSweeps=0)
    b, solo = make_pair(SIZES, knobs)
    b.BlockSweeps, b.BlockCols = shape
    b.HandoffLag = 2 + (shape[0] + shape[1]) % 2  # a block late or halfway
    for step in range(3):
        b.update(C.DT)
        for s in solo:
            s.update(C.DT)
        for e, (m, s) in enumerate(zip(b.members, solo)):
            same(m, s, "shape %s step %d member %d %s" % (shape, step, e, SIZES[e]))
    b.close()
    for s in solo:
        s.close()


def test_batch_multigrid_members_match_solo_and_oracle():
    """UseMultigrid on a batch: every member takes the V-cycle in its own update with the
    batch's knobs, bit-identical to stepping it alone and to the oracle; switching back to the
    merged single-grid solve afterwards follows the members' traded buffers."""
    sizes = [(48, 40), (130, 70), (64, 64), (33, 77)]
    knobs = dict(UseBFECC=True, PressureIters=4, EarlyExit=True)
    b, solo = make_pair(sizes, knobs)
    o = OracleFluid(C.DENSITY, 33, 77, C.H)
    scene(3, o)
    for k, v in dict(BASE_KNOBS, **knobs).items():
        if k not in ("BlockSweeps", "HandoffLag"):
            setattr(o, k, v)
    for f in [b, o] + solo:
        f.UseMultigrid, f.MultigridLevels = True, 2
    for step in range(3):
        b.update(C.DT)
        for s in solo:
            s.update(C.DT)
        o.update(C.DT)
        for e, (m, s) in enumerate(zip(b.members, solo)):
            same(m, s, "multigrid step %d member %d %s" % (step, e, sizes[e]))
        same(b.members[3], o, "multigrid step %d oracle" % step)
    for f in [b] + solo:
        f.UseMultigrid, f.PressureIters = False, 20
    for step in range(2):
        b.update(C.DT)
        for s in solo:
            s.update(C.DT)
        for e, (m, s) in enumerate(zip(b.members, solo)):
            same(m, s, "single-grid step %d member %d" % (step, e))
    b.close()
    for s in solo:
        s.close()


def test_batch_member_vs_oracle():
    """One ensemble member against the oracle directly (no solo Fluid in between)."""
    sizes = [(90, 50), (33, 77)]
    b = FluidBatch(C.DENSITY, sizes, C.H)
    o = OracleFluid(C.DENSITY, 33, 77, C.H)
    for e, m in enumerate(b.members):
        scene(e + 2, m)
    scene(3, o)
    knobs = dict(BASE_KNOBS, **KNOB_SETS["bfecc_conf"])
    set_knobs(b, knobs)
    for k, v in knobs.items():
        if k != "BlockSweeps":
            setattr(o, k, v)
    for step in range(3):
        b.update(C.DT)
        o.update(C.DT)
        same(b.members[1], o, "step %d" % step)
    b.close()


@pytest.mark.parametrize("kset", ["bfecc_conf", "sl_visc"])
def test_graph_replays_track_changes(kset):
    """Replays of the captured step stay identical to solo stepping while edits, knob
    changes, new sources and a member's own Simulate happen between them. With
    semi-Lagrangian smoke every step trades the M and M2 roles of all members."""
    sizes = SIZES[:5]
    knobs = KNOB_SETS[kset]
    b, solo = make_pair(sizes, knobs)

    def both(fn):
        fn(b.members)
        fn(solo)

    def check(where):
        for e, (m, s) in enumerate(zip(b.members, solo)):
            same(m, s, "%s member %d" % (where, e))

    def step():
        b.update(C.DT)
        for s in solo:
            s.update(C.DT)

    for _ in range(3):
        step()
    check("replays")
    # edits (flags change) between replays
    both(lambda fs: (fs[1].SetSolid(5, 5, False), fs[2].SetVelocity(3, 4, 0.25, -0.5), fs[0].AddSmoke(2, 2, 1.0)))
    step()
    check("after edits")
    # a knob change recaptures
    for o in [b] + solo + b.members:
        o.PressureIters = 30
        o.Confinement = 0.1
    step()
    check("after knob change")
    # a new source changes the source program's launch shape
    both(lambda fs: fs[3].add_source(4, 6, 0.5, 0.25))
    step()
    check("after new source")
    # a member stepped on its own swaps its buffers; the next batched step follows
    b.members[4].update(C.DT)
    solo[4].update(C.DT)
    step()
    step()
    check("after member's own step")
    b.close()


def test_autotune_keeps_state():
    sizes = [(120, 80), (64, 200)]
    b, solo = make_pair(sizes, KNOB_SETS["bfecc_conf"])
    b.update(C.DT)
    for s in solo:
        s.update(C.DT)
    before = [m.state() for m in b.members]
    tk, ms = b.autotune(C.DT)
    # one of the candidate shapes (2-sweep tiles are tried only on grids this small)
    assert tk in (2, 4, 8, 16) and b.BlockCols in (4, 8, 16) and b.HandoffLag in (1, 2, 3) and ms > 0
    for m, st in zip(b.members, before):
        for k, v in st.items():
            np.testing.assert_array_equal(m.get(k).view(np.uint32), v.view(np.uint32))
    # after tuning the batch keeps stepping identically to the solo runs at the tuned depth
    for s in solo:
        s.BlockSweeps = tk
        s.update(C.DT)
    b.update(C.DT)
    for m, s in zip(b.members, solo):
        same(m, s, "after autotune")
    # single-context autotuner
    f = solo[0]
    st = f.state()
    tk1, ms1 = f.autotune(C.DT)
    assert tk1 in (2, 4, 8, 16) and f.BlockCols in (4, 8, 16) and f.HandoffLag in (1, 2, 3) and ms1 > 0
    for k, v in st.items():
        np.testing.assert_array_equal(f.get(k).view(np.uint32), v.view(np.uint32))
    b.close()


def test_batch_rejects_bad_input():
    from fluid_amd import FluidError
    with pytest.raises(FluidError):
        FluidBatch(C.DENSITY, [(4, 4), (0, 3)], C.H)
    b = FluidBatch(C.DENSITY, [(8, 8)], C.H)
    b.PressureIters = 5000
    with pytest.raises(FluidError):
        b.update(C.DT)
    b.close()


def test_gaussian_impulse_matches_oracle():
    g, o = Fluid(C.DENSITY, 70, 50, C.H), OracleFluid(C.DENSITY, 70, 50, C.H)
    for f in (g, o):
        C.open_all_cells(f)
        f.SetSolid(33, 25, False)
        f.ApplyForceGaussian(30, 25, 3.5, -1.25, 8.0)
        f.ApplyForceGaussian(2, 48, -0.5, 4.0, 2.5)
    same(g, o, "impulses", ("U", "V", "S"))


def test_ensemble_workload_small():
    """The ensemble workload (presets, wall strokes, Gaussian impulses every 10 steps,
    60 sweeps, SL) on 12 members of sides 64..256: batched == solo, and three members
    against the oracle."""
    specs = C.ensemble_specs(n=12, seed=9, sides=tuple(range(64, 257, 32)))
    b = FluidBatch(C.DENSITY, [(s["w"], s["h"]) for s in specs], C.H)
    solo = [Fluid(C.DENSITY, s["w"], s["h"], C.H) for s in specs]
    picks = (0, 5, 11)
    ora = {e: OracleFluid(C.DENSITY, specs[e]["w"], specs[e]["h"], C.H, threads=8) for e in picks}
    for e, s in enumerate(specs):
        for f in (b.members[e], solo[e]) + ((ora[e],) if e in ora else ()):
            C.ensemble_member_setup(f, s)
    C.ensemble_knobs(b)
    for step in range(12):
        for (e, cx, cy, fx, fy, sg) in C.ensemble_impulses(step, specs):
            b.members[e].ApplyForceGaussian(cx, cy, fx, fy, sg)
            solo[e].ApplyForceGaussian(cx, cy, fx, fy, sg)
            if e in ora:
                ora[e].ApplyForceGaussian(cx, cy, fx, fy, sg)
        b.update(C.DT)
        for f in solo + list(ora.values()):
            f.update(C.DT)
        if step in (0, 1, 10, 11):
            for e in range(len(specs)):
                same(b.members[e], solo[e], "step %d member %d" % (step, e))
            for e, o in ora.items():
                same(b.members[e], o, "oracle step %d member %d" % (step, e))
    b.close()
