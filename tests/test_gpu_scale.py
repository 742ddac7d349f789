"""GPU parity at scale: grids that span many temporal-blocked tiles in both directions,
the Karman workload at its full size, and the 4096^2 cavity (bit-exact at full grid size
with 40 sweeps; with 200 sweeps through size-independent properties)."""
import numpy as np
import pytest

from fluid_amd import Fluid
from fluid_amd import configs as C
from oracle.pyoracle import OracleFluid

pytestmark = pytest.mark.gpu

FIELDS = ("U", "V", "newU", "newV", "p", "S", "M", "newM")


def same(g, o, where, fields=FIELDS):
    for k in fields:
        a, b = g.get(k), o.get(k)
        bad = a.view(np.uint32) != b.view(np.uint32)
        if bad.any():
            i, j = np.argwhere(bad)[0]
            raise AssertionError("%s: %s differs at %d cells, first (%d,%d) gpu %r oracle %r" %
                                 (where, k, int(bad.sum()), i, j, a[i, j], b[i, j]))


# every tile shape the solver and its autotuner can pick: sweeps per tile x columns per block
SHAPES = [(tk, blk) for tk in (2, 4, 8, 16) for blk in (4, 8, 16)]
SHAPE_IDS = ["tk%d_blk%d" % s for s in SHAPES]
_ORACLE = {}


def many_tiles_reference(early):
    """Initial state and the oracle's state after each of 3 steps: 700x300, 100 sweeps (15 to
    50 blocks of sweeps, so consumers of empty G producers poll their corner), BFECC."""
    if early not in _ORACLE:
        w, h, n = 700, 300, 100
        rng = np.random.default_rng(21)
        st = {}
        S = (rng.random((w + 2, h + 2)) > 0.05).astype(np.float32)
        S[0, :] = S[:, 0] = S[:, -1] = 0
        st["S"] = S
        for k in ("U", "V", "newU", "newV"):
            st[k] = (0.5 * rng.standard_normal((w + 2, h + 2))).astype(np.float32)
        st["M"] = rng.random((w + 2, h + 2)).astype(np.float32)
        o = OracleFluid(1000.0, w, h, 0.01)
        o.load_state(st)
        o.PressureIters, o.EarlyExit, o.UseBFECC = n, early, True
        outs = []
        for _ in range(3):
            o.Simulate(0.1)
            outs.append(o.state())
        _ORACLE[early] = (st, outs)
    return _ORACLE[early]


@pytest.mark.parametrize("shape", SHAPES, ids=SHAPE_IDS)
@pytest.mark.parametrize("early", [False, True])
@pytest.mark.parametrize("lag", [1, 2, 3], ids=["on_time", "one_block_late", "halfway"])
def test_many_tiles_bit_exact(shape, early, lag):
    st, outs = many_tiles_reference(early)
    g = Fluid(1000.0, 700, 300, 0.01)
    g.load_state(st)
    g.PressureIters, g.EarlyExit, g.UseBFECC = 100, early, True
    g.BlockSweeps, g.BlockCols = shape
    g.HandoffLag = lag  # both ways of publishing the tile handoffs
    for step in range(3):
        g.Simulate(0.1)
        for k in FIELDS:
            a, b = g.get(k), outs[step][k]
            bad = a.view(np.uint32) != b.view(np.uint32)
            assert not bad.any(), "step %d: %s differs at %d cells" % (step, k, int(bad.sum()))
    g.close()


def test_gs_layout_reports_the_tiling():
    """fl_get_gs_layout: an error before the first temporal-blocked solve, then the tiling that
    solve used (150x90, 24 sweeps, 8 sweeps x 4 columns, halfway publication)."""
    from fluid_amd._lib import FluidError
    g = Fluid(1000.0, 150, 90, 0.01)
    with pytest.raises(FluidError, match="no temporal-blocked solve"):
        g.gs_layout()
    g.PressureIters, g.EarlyExit = 24, False
    g.BlockSweeps, g.BlockCols, g.HandoffLag = 8, 4, 3
    g.Simulate(0.1)
    L = g.gs_layout()
    # rows i' = i + k span 0..151+23 over 64-row tiles; 92 + 63 + 7 column steps rounded up to blocks
    assert (L["sweeps_per_tile"], L["cols_per_block"], L["handoff_lag"]) == (8, 4, 3)
    assert (L["row_tiles"], L["sweep_blocks"], L["steps_per_tile"]) == (3, 3, 164)
    assert 1 <= L["workgroups"] <= L["tiles"] <= 9
    g.close()


@pytest.mark.parametrize("shape", SHAPES, ids=SHAPE_IDS)
def test_small_grids_every_shape(shape):
    """Grids of one tile, one strip and a few tiles (the critical-path-bound sizes the tuner
    tries shallow tiles on), every handoff publication, bit-exact with the oracle."""
    for (w, h, n) in ((1, 1, 6), (37, 23, 5), (150, 90, 24), (40, 300, 70)):
        rng = np.random.default_rng(w * 1000 + h)
        st = {"S": (rng.random((w + 2, h + 2)) > 0.1).astype(np.float32)}
        for k in ("U", "V", "newU", "newV"):
            st[k] = (0.5 * rng.standard_normal((w + 2, h + 2))).astype(np.float32)
        o = OracleFluid(1000.0, w, h, 0.01)
        o.load_state(st)
        o.PressureIters, o.EarlyExit = n, False
        o.Simulate(0.1)
        want = o.state()
        for lag in (1, 2, 3):
            g = Fluid(1000.0, w, h, 0.01)
            g.load_state(st)
            g.PressureIters, g.EarlyExit = n, False
            g.BlockSweeps, g.BlockCols = shape
            g.HandoffLag = lag
            g.Simulate(0.1)
            for k in FIELDS:
                a, b = g.get(k), want[k]
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (w, h, n, lag, k)
            g.close()


def test_karman_full_size():
    """The Karman street at 2048x1024: disc r=64, inflow u=1, BFECC, confinement 0.3, 100 sweeps."""
    g = C.karman_scene(lambda *a: Fluid(*a))
    o = C.karman_scene(lambda *a: OracleFluid(*a, threads=8))
    same(g, o, "setup")
    for step in range(2):
        g.update(C.DT)
        o.update(C.DT)
        same(g, o, "step %d" % step)


def test_cavity_full_grid_bit_exact():
    """The 4096^2 cavity grid (lid, seeded discs, BFECC), one step with 40 sweeps."""
    g = C.cavity_scene(lambda *a: Fluid(*a), iters=40)
    o = C.cavity_scene(lambda *a: OracleFluid(*a, threads=8), iters=40)
    same(g, o, "setup", ("S", "U", "V"))
    g.update(C.DT)
    o.update(C.DT)
    same(g, o, "step")


def test_cavity_full_size_properties():
    """The 4096^2 cavity with 200 sweeps: the result is bit-identical from run to run
    (the tile handoffs are race-free), the sweeps cut the divergence, the solve keeps all
    200 sweeps, and nothing becomes NaN."""
    g = C.cavity_scene(lambda *a: Fluid(*a))
    for _ in range(3):
        g.update(C.DT)
    st = g.state()
    div0 = None
    outs = []
    for rep in range(2):
        g.load_state(st)
        if rep == 0:
            div0 = g.MaxDivergence()
        g.update(C.DT)
        outs.append({k: g.get(k) for k in ("U", "V", "p", "M")})
        s = g.stats()
        assert s["sweeps_run"] == 200 and not s["invalid_sample"]
    for k in outs[0]:
        assert np.array_equal(outs[0][k].view(np.uint32), outs[1][k].view(np.uint32)), k
        assert np.isfinite(outs[0][k]).all(), k
    # the projection alone reduces the divergence of the post-source state
    g.load_state(st)
    g.run_pass("clearP")
    g.run_pass("makeIncompressible", C.DT, 200)
    assert g.MaxDivergence() < 0.5 * div0


def test_full_size_early_exit_rollback():
    """A quiescent 4096^2 field converges in sweep 0: the device rolls back and keeps 1 sweep."""
    g = Fluid(1000.0, 4096, 4096, 0.01)
    C.open_all_cells(g)
    g.scene_walls(True, True, True, True)
    g.apply_wall_settings()
    g.PressureIters = 200
    g.Simulate(0.1)
    s = g.stats()
    assert s["sweeps_run"] == 1 and s["reran"]
    assert not g.get("U").any() and not g.get("p").any()
