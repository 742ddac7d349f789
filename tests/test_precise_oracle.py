"""PreciseAdvection, an addition with no reference counterpart: every back-trace is taken in
cell units, X = i + (s*u)*h1 with s = -dt or +dt, instead of at the reference's absolute
position i*h + offset + s*u, which float32 rounds at ulp(N*h) (sampleField, fluid.go:357-398).
These tests pin the oracle's restatement of it:

- where every position is exact in float32 (h a power of two, dt*u on a dyadic grid) the
  precise and the reference sampling are the same function bit for bit, clamps at all four
  edges included (displacements beyond the grid, positions exactly on the clamp bounds);
- in FP64 the two are the same function up to rounding, over whole steps with BFECC and
  over the adjoint: a wrong half-cell offset, sign or clamp bound differs by O(1);
- in float32 the precise step is within 1e-5 of the FP64 step from the same state, where the
  reference arithmetic is several times further off."""
import numpy as np
import pytest

from fluid_amd import configs as C
from oracle.pyoracle import OracleFluid


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def exact_position_state(nx, ny, h, dt, seed):
    """Fields whose back-traces all land on exact float32 positions: displacements are
    multiples of 1/4 cell up to a grid width and a half past either edge, so many land
    exactly on the clamp bounds 1, N, 1/2 and N - 1/2."""
    rng = np.random.default_rng(seed)
    shape = (nx, ny)
    span = max(nx, ny) + 6
    # U, V as displacement/dt*h with displacement in quarter cells; face averages of four
    # such values are sixteenths of a cell, still exact after *dt and *h1
    U = (rng.integers(-4 * span, 4 * span + 1, shape) / 4.0 * h / dt).astype(np.float32)
    V = (rng.integers(-4 * span, 4 * span + 1, shape) / 4.0 * h / dt).astype(np.float32)
    S = (rng.random(shape) > 0.12).astype(np.float32)
    S[0, :] = S[-1, :] = S[:, 0] = S[:, -1] = 0.0
    M = rng.random(shape).astype(np.float32)
    return dict(U=U, V=V, S=S, M=M, newU=rng.random(shape).astype(np.float32),
                newV=rng.random(shape).astype(np.float32), newM=rng.random(shape).astype(np.float32))


@pytest.mark.parametrize("h,dt", [(1.0, 0.25), (0.5, 0.125)], ids=["h1", "h0.5"])
@pytest.mark.parametrize("pass_", ["advectVelocity", "advectSmoke"])
def test_precise_equals_reference_where_positions_are_exact(h, dt, pass_):
    st = exact_position_state(26, 19, h, dt, seed=7)
    out = []
    for precise in (False, True):
        o = OracleFluid(1000.0, 24, 17, h)
        o.load_state(st)
        o.PreciseAdvection = precise
        o.run_pass(pass_, dt)
        assert not o.invalid_index()
        out.append({k: o.get(k) for k in ("U", "V", "M", "newU", "newV", "newM")})
    for k in out[0]:
        np.testing.assert_array_equal(out[0][k], out[1][k], err_msg=k)
    # the state really exercised the clamps: some faces moved, and displacements exceed the grid
    assert np.abs(st["U"]).max() * dt / h > 26


def test_exact_positions_hit_every_clamp_bound():
    """The displacement grid above lands back-traces exactly on each clamp bound (X = 1,
    X = N for the unshifted axes, X = 1/2, N - 1/2 for the half-shifted ones) and beyond."""
    nx, ny, h, dt = 26, 19, 1.0, 0.25
    st = exact_position_state(nx, ny, h, dt, seed=7)
    i = np.arange(nx)[:, None].astype(np.float64)
    X = i - dt * st["U"].astype(np.float64) / h  # U faces, x axis: bounds [1, NX]
    assert (X == 1).any() and (X == nx).any() and (X < 1).any() and (X > nx).any()
    j = np.arange(ny)[None, :].astype(np.float64)
    Y = j - dt * st["V"].astype(np.float64) / h  # V faces, y axis: bounds [1, NY]
    assert (Y == 1).any() and (Y == ny).any()


def fp64_step_errors(scene, steps, **knobs):
    """Max over a trajectory of the FP64 reference model of the one-step difference between
    the reference sampling and PreciseAdvection, each step taken from the identical state.
    (Free-running trajectories separate: BFECC amplifies a 1e-15 difference 3-5x per step.)"""
    pair = []
    for precise in (False, True):
        o = scene(lambda *a: OracleFluid(*a, prec=64))
        for k, v in knobs.items():
            setattr(o, k, v)
        o.EarlyExit = False
        o.PreciseAdvection = precise
        pair.append(o)
    a, b = pair
    worst = {k: 0.0 for k in ("U", "V", "p", "M")}
    for _ in range(steps):
        b.load_state(a.state())
        a.update(C.DT)
        b.update(C.DT)
        for k in worst:
            worst[k] = max(worst[k], rel(b.get(k), a.get(k)))
    assert np.abs(a.get("U")).max() > 0.1  # the flow developed
    return worst


@pytest.mark.parametrize("bfecc", [False, True], ids=["sl", "bfecc"])
def test_precise_is_the_reference_function_in_fp64_jet(bfecc):
    for k, e in fp64_step_errors(lambda mk: C.jet_scene(mk, 40, 32, 20, jet_rows=8), 12, UseBFECC=bfecc).items():
        assert e < 1e-12, (k, e)


def test_precise_is_the_reference_function_in_fp64_cavity():
    for k, e in fp64_step_errors(lambda mk: C.cavity_scene(mk, 48, 40, 30, lid_u=3.0), 10).items():
        assert e < 1e-12, (k, e)


def test_precise_is_the_reference_function_in_fp64_open_edge():
    """Open right edge with an obstacle and an inflow column: back-traces leave the grid
    through the clamps on the open side."""
    def scene(mk):
        f = C.karman_scene(mk, 64, 32, 20, radius=5, centre=(16, 16), band=6)
        f.Confinement = 0.0
        return f
    for k, e in fp64_step_errors(scene, 10).items():
        assert e < 1e-12, (k, e)


def test_precise_float32_step_is_within_1e5_of_fp64():
    """One cavity step (256^2, BFECC, 200 sweeps) from a developed state: the float32 step
    with the reference's arithmetic against FP64, and with PreciseAdvection. The FP64 step
    follows float32's sweep count and turbulence gate (both are decided before advection)."""
    mk32 = lambda *a: OracleFluid(*a, threads=8)
    g = C.cavity_scene(mk32, 256, 256, 200)
    for _ in range(6):
        g.update(C.DT)
    st = g.state()
    runs = {}
    for name, prec, precise in (("ref32", 32, False), ("precise32", 32, True), ("fp64", 64, False)):
        o = C.cavity_scene(lambda *a: OracleFluid(*a, prec=prec, threads=8), 256, 256, 200)
        o.load_state(st)
        o.PreciseAdvection = precise
        if prec == 64:
            o.set_forced_sweeps(runs["ref32"].last_sweeps())
            o.set_turb_gate(runs["ref32"].turb_gate())
        o.update(C.DT)
        runs[name] = o
    for k in ("U", "V", "M"):
        e_ref = rel(runs["ref32"].get(k), runs["fp64"].get(k))
        e_pre = rel(runs["precise32"].get(k), runs["fp64"].get(k))
        assert e_pre < 1e-5, (k, e_pre)
        assert e_ref > 3 * e_pre, (k, e_ref, e_pre)
    assert rel(runs["precise32"].get("p"), runs["fp64"].get("p")) < 1e-5


def adjoint_scene(precise, bfecc=True):
    o = OracleFluid(C.DENSITY, 20, 16, C.H, prec=64)
    C.jet_scene(lambda *a: o, 20, 16, 6, jet_u=1.0, jet_rows=4, walls=(True, False, True, False))
    o.SetCircularObstacle(10, 8, 2)
    o.UseBFECC, o.EarlyExit, o.PreciseAdvection = bfecc, False, precise
    return o


def rollout_inputs(o, seed=0):
    rng = np.random.default_rng(seed)
    shape = (o.NumX, o.NumY)
    return 0.05 * rng.standard_normal(shape), 0.05 * rng.standard_normal(shape), rng.random(shape)


@pytest.mark.parametrize("bfecc", [False, True], ids=["sl", "bfecc"])
def test_precise_adjoint_matches_finite_differences(bfecc):
    o = adjoint_scene(True, bfecc)
    Fu, Fv, T = rollout_inputs(o)
    loss, gu, gv = o.rollout(3, C.DT, Fu, Fv, T)
    rng = np.random.default_rng(1)
    for _ in range(4):
        du, dv = rng.standard_normal(Fu.shape), rng.standard_normal(Fu.shape)
        eps = 1e-6
        fd = (o.rollout(3, C.DT, Fu + eps * du, Fv + eps * dv, T, grad=False) -
              o.rollout(3, C.DT, Fu - eps * du, Fv - eps * dv, T, grad=False)) / (2 * eps)
        ad = (gu * du).sum() + (gv * dv).sum()
        assert abs(fd - ad) / abs(fd) < 1e-5


def test_precise_adjoint_is_the_reference_adjoint_in_fp64():
    """Same loss and gradient as the reference sampling's adjoint, to rounding: the cell-unit
    geometry changes neither the corners, the weights nor the derivative masks."""
    res = []
    for precise in (False, True):
        o = adjoint_scene(precise)
        res.append(o.rollout(3, C.DT, *rollout_inputs(o)))
    (l0, gu0, gv0), (l1, gu1, gv1) = res
    assert abs(l1 - l0) <= 1e-12 * abs(l0)
    assert rel(gu1, gu0) < 1e-10 and rel(gv1, gv0) < 1e-10
