"""Pins the CPU oracle (oracle/fluid_oracle.cpp) before it is trusted as the checker.

Each check is chosen so that a plausible restatement mistake (a dropped term, a wrong
sign, a wrong face index, a transposed operand) fails it: the reference's own
known-answer tests (restated from pkg/fluid/*_test.go), closed forms, invariants of
the projection, and agreement of the float32 and float64 instantiations.
"""
import copy
import math
import struct

import numpy as np
import pytest

from known_answers import load_cases, run_case
from oracle.pyoracle import OracleFluid, go_sin, go_sin_many, noise_table

F32 = np.float32


def fluid(w, h, density=1.0, hh=1.0, prec=32):
    f = OracleFluid(density, w, h, hh, prec=prec)
    f.set("S", np.ones((f.NumX, f.NumY)))
    return f


# ---- the Go/Cephes sine (third-party arithmetic of addTurbulence) --------------------------

def test_sin_constants_match_their_bit_patterns():
    # Decimal constants of Go's sin.go paired with the IEEE bit patterns published beside them.
    pairs = [(7.85398125648498535156e-1, 0x3fe921fb40000000), (3.77489470793079817668e-8, 0x3e64442d00000000),
             (2.69515142907905952645e-15, 0x3ce8469898cc5170), (1.58962301576546568060e-10, 0x3de5d8fd1fd19ccd),
             (-1.66666666666666307295e-1, 0xbfc5555555555548), (4.16666666666665929218e-2, 0x3fa555555555554b)]
    for v, bits in pairs:
        assert struct.unpack("<Q", struct.pack("<d", v))[0] == bits


def test_go_sin_against_libm():
    x = np.concatenate([np.linspace(-20, 20, 20001), np.linspace(0, 17000, 50001)])
    g = go_sin_many(x)
    ref = np.array([math.sin(v) for v in x])
    ok = np.abs(ref) > 1e-300
    ulp = np.abs(g[ok] - ref[ok]) / np.spacing(np.abs(ref[ok]))
    assert ulp.max() <= 1.0
    assert go_sin(0.0) == 0.0 and math.isnan(go_sin(float("nan")))
    assert go_sin(-1.25) == -go_sin(1.25)


def test_turbulence_noise_table_is_implementation_independent():
    """float32(sin(float32(k)*0.01)) for every seed of a 4096^2 grid agrees with libm and
    no seed lies within 4 double ulps of a float32 rounding boundary, so any sine within
    a few ulps (Go's, libm's, the device's) yields the same float32 table."""
    n = 263 * 4098 + 157 * 4098 + 1
    t = noise_table(n)
    seed = (np.arange(n).astype(F32) * F32(0.01)).astype(np.float64)
    s = np.sin(seed)
    assert int((s.astype(F32) != t).sum()) == 0
    # distance of s to the nearest float32 midpoint, in double ulps
    f = s.astype(F32).astype(np.float64)
    up = np.nextafter(s.astype(F32), F32(np.inf)).astype(np.float64)
    dn = np.nextafter(s.astype(F32), F32(-np.inf)).astype(np.float64)
    mid = np.where(s >= f, (f + up) / 2, (f + dn) / 2)
    dist = np.abs(s - mid) / np.spacing(np.abs(s) + 1e-300)
    assert dist[np.abs(s) > 1e-30].min() > 4.0


# ---- restated reference tests with exact answers ---------------------------------------------

KNOWN = load_cases()


@pytest.mark.parametrize("case", KNOWN, ids=[c["name"] for c in KNOWN])
def test_reference_known_answers(case):
    """The fixture tests/golden/reference_known_answers.json: each reference test's own
    assertions plus the closed forms of the same state."""
    n, fails = run_case(lambda *a: OracleFluid(*a), case)
    assert n > 0 and not fails, fails


def test_known_answer_runner_reports_a_wrong_answer():
    """The runner is not vacuous: a case whose expected value is off by one float32 ulp, or
    whose expected view value has the wrong sign, is reported."""
    case = copy.deepcopy(next(c for c in KNOWN if c["name"] == "TestApplyForceRadius"))
    chk = case["exact"][0]
    chk[5] = float(np.nextafter(F32(chk[5]), F32(1e9)))
    assert len(run_case(lambda *a: OracleFluid(*a), case)[1]) == 1
    case = copy.deepcopy(next(c for c in KNOWN if c["name"] == "TestVorticityField"))
    case["exact"][0][5] = -2.0
    assert len(run_case(lambda *a: OracleFluid(*a), case)[1]) == 1


def test_copy_border_exact():
    # TestCopyBorder (fluid_test.go:8-35)
    f = OracleFluid(1.0, 3, 3, 1.0)
    src = np.arange(1, 26, dtype=np.float32).reshape(5, 5)
    f.set("U", src)
    f.set("V", src * 2)
    f.run_pass("copyBorderUV")
    nu, nv = f.get("newU"), f.get("newV")
    ring = np.ones((5, 5), bool)
    ring[1:4, 1:4] = False
    assert np.array_equal(nu[ring], src[ring]) and np.array_equal(nv[ring], 2 * src[ring])
    assert not nu[~ring].any()


def test_apply_force_exact():
    # TestApplyForce (fluid_test.go:1134-1160)
    f = fluid(10, 10)
    f.ApplyForce(5, 5, 3.0, -2.0)
    U, V = f.get("U"), f.get("V")
    assert U[5, 5] == 3.0 and V[5, 5] == -2.0
    S = f.get("S")
    S[6, 6] = 0
    f.set("S", S)
    f.ApplyForce(6, 6, 10.0, 10.0)
    f.ApplyForce(0, 0, 1.0, 1.0)
    U, V = f.get("U"), f.get("V")
    assert U[6, 6] == 0 and V[6, 6] == 0 and U[0, 0] == 0


def test_apply_force_radius_closed_form():
    # TestApplyForceRadius (fluid_test.go:1162-1190): weight exp(-3 d^2/r^2) inside the disc.
    f = fluid(20, 20)
    f.ApplyForceRadius(10, 10, 5.0, 0.0, 3)
    U = f.get("U")
    assert U[10, 10] == F32(5.0)
    assert U[13, 10] == F32(5.0) * F32(math.exp(-3.0))
    assert U[12, 11] == F32(5.0) * F32(math.exp(float(F32(-3.0) * F32(5.0) / F32(9.0))))
    assert U[14, 10] == 0 and U[13, 13] == 0


def test_apply_force_gaussian_closed_form():
    # The ensemble's impulse: ApplyForce weighted by exp(-d^2/(2 sigma^2)), d <= 3 sigma.
    f = fluid(70, 70)
    S = f.get("S")
    S[30, 40] = 0
    f.set("S", S)
    f.ApplyForceGaussian(30, 30, 2.0, -1.0, 4.0)
    U, V = f.get("U"), f.get("V")
    assert U[30, 30] == F32(2.0) and V[30, 30] == F32(-1.0)              # centre weight exp(0)
    assert U[34, 30] == F32(2.0) * F32(math.exp(-0.5))                    # d = sigma
    assert U[30, 18] == F32(2.0) * F32(math.exp(-4.5))                    # d = 3 sigma, kept
    assert U[42, 31] == 0 and U[30, 17] == 0                              # d > 3 sigma
    assert U[30, 40] == 0 and V[30, 40] == 0                              # solid cell skipped
    assert np.array_equal(U[30:55, 30], U[30:5:-1, 30])                   # symmetric
    assert (U != 0).sum() == sum(1 for a in range(-12, 13) for b in range(-12, 13)
                                 if a * a + b * b <= 144) - 1            # the disc minus the solid cell
    # clipped at the interior like ApplyForce
    g = fluid(20, 20)
    g.ApplyForceGaussian(1, 1, 1.0, 1.0, 3.0)
    U = g.get("U")
    assert U[0, :].max() == 0 and U[:, 0].max() == 0 and U[1, 1] == 1.0


def test_set_circular_obstacle_and_set_solid_faces():
    # TestSetCircularObstacle (fluid_test.go:1427-1447) + SetSolid face zeroing (walls.go:19-48)
    f = fluid(20, 20)
    for k in ("U", "V", "newU", "newV"):
        f.set(k, np.full((22, 22), 7.0))
    f.SetCircularObstacle(10, 10, 3)
    assert f.IsSolid(10, 10) and f.IsSolid(12, 10) and not f.IsSolid(14, 10)
    S = f.get("S")
    ii, jj = np.meshgrid(np.arange(22), np.arange(22), indexing="ij")
    disc = (ii - 10) ** 2 + (jj - 10) ** 2 <= 9
    assert np.array_equal(S == 0, disc)
    for k in ("U", "newU"):
        a = f.get(k)
        zero = disc | np.roll(disc, 1, axis=0)  # faces i and i+1 of each solid cell
        assert np.array_equal(a == 0, zero)
    for k in ("V", "newV"):
        a = f.get(k)
        zero = disc | np.roll(disc, 1, axis=1)
        assert np.array_equal(a == 0, zero)


def test_walls_panics():
    f = fluid(4, 4)
    with pytest.raises(IndexError, match="invalid x-index: 6"):
        f.SetSolid(6, 0, True)
    with pytest.raises(IndexError, match="invalid y-index: -1"):
        f.AddSmoke(0, -1, 1.0)


def test_vorticity_and_magnitude_closed_forms():
    # TestVorticityField / TestVelocityMagnitudeField (fluid_test.go:1381-1425)
    f = fluid(10, 10)
    U = np.zeros((12, 12), np.float32)
    V = np.zeros((12, 12), np.float32)
    U[5, 4], U[5, 6], V[4, 5], V[6, 5] = 1, -1, -1, 1
    f.set("U", U)
    f.set("V", V)
    assert f.Vorticity().Value(5, 5)[0] == 2.0
    g = fluid(10, 10)
    U = np.zeros((12, 12), np.float32)
    V = np.zeros((12, 12), np.float32)
    U[5, 5] = U[6, 5] = 3.0
    V[5, 5] = V[5, 6] = 4.0
    g.set("U", U)
    g.set("V", V)
    vm = g.VelocityMagnitude()
    assert vm.Value(5, 5)[0] == 5.0 and vm.MaxValue == 5.0


def test_handle_borders_exact():
    # TestBoundaryConditions part 1 (fluid_test.go:287-329) plus the extrapolation branch.
    f = fluid(6, 6)
    S = f.get("S")
    S[3, 3] = 0
    S[:, 0] = 0  # bottom wall
    f.set("S", S)
    U = np.ones((8, 8), np.float32)
    U[:, 1] = 3.0
    U[:, 2] = 5.0
    f.set("U", U)
    f.run_pass("handleBorders")
    U = f.get("U")
    assert (U[:, 0] == 0).all()
    # top row: S open on both sides -> 2*U[i,NY-2] - U[i,NY-3] for interior i
    assert (U[1:7, 7] == 1.0).all()


def test_confinement_changes_vortex():
    # TestApplyVorticityConfinement (confinement_test.go:6-28)
    f = OracleFluid(1, 4, 4, 1)
    for i in range(6):
        for j in range(6):
            f.SetSolid(i, j, False)
    f.SetVelocity(3, 2, 1, 0)
    f.SetVelocity(2, 3, -1, 0)
    f.SetVelocity(2, 2, 0, 1)
    f.SetVelocity(3, 3, 0, -1)
    f.Confinement = 5
    u0, v0 = f.get("U")[2, 2], f.get("V")[2, 2]
    f.run_pass("applyVorticityConfinement", 1.0)
    assert f.get("U")[2, 2] != u0 or f.get("V")[2, 2] != v0


# ---- pressure projection: closed forms and invariants ----------------------------------------

def test_single_cell_sweep_closed_form():
    """One isolated fluid cell: q = (-div/s)*omega*damping; each fluid-side face moves by q."""
    f = OracleFluid(2.0, 1, 1, 0.5)
    S = np.ones((3, 3), np.float32)
    S[0, 1] = 0  # left neighbour solid -> s = 3
    f.set("S", S)
    U = np.zeros((3, 3), np.float32)
    V = np.zeros((3, 3), np.float32)
    U[1, 1], U[2, 1], V[1, 1], V[1, 2] = 0.3, 1.1, -0.2, 0.7
    f.set("U", U)
    f.set("V", V)
    f.Relaxation = 1.9
    f.PressureDamping = 0.5
    f.run_pass("makeIncompressible", 0.25, 1)
    div = (F32(1.1) - F32(0.3)) + F32(0.7) - F32(-0.2)
    q = (-div / F32(3)) * F32(1.9) * F32(0.5)
    cp = F32(2.0) * F32(0.5) / F32(0.25)
    U, V, p = f.get("U"), f.get("V"), f.get("p")
    assert U[1, 1] == F32(0.3)            # solid side: weight 0
    assert U[2, 1] == F32(1.1) + q
    assert V[1, 1] == F32(-0.2) - q
    assert V[1, 2] == F32(0.7) + q
    assert p[1, 1] == cp * q


def test_gauss_seidel_unit_relaxation_zeroes_each_cell_divergence():
    """With omega = 1 each update makes its own cell divergence-free: after one sweep the
    last cell of the lexicographic order has (numerically) zero divergence, while an
    earlier one does not. Pins the face indices, signs and the serial order."""
    rng = np.random.default_rng(0)
    f = OracleFluid(1.0, 6, 5, 1.0)
    S = np.ones((8, 7), np.float32)
    S[0, :] = S[-1, :] = S[:, 0] = S[:, -1] = 0
    f.set("S", S)
    f.set("U", rng.standard_normal((8, 7)).astype(np.float32))
    f.set("V", rng.standard_normal((8, 7)).astype(np.float32))
    f.Relaxation = 1.0
    f.run_pass("makeIncompressible", 1.0, 1)
    U, V = f.get("U").astype(np.float64), f.get("V").astype(np.float64)
    div = lambda i, j: U[i + 1, j] - U[i, j] + V[i, j + 1] - V[i, j]
    assert abs(div(6, 5)) < 1e-6
    assert abs(div(1, 1)) > 1e-3


def test_projection_applies_a_pressure_gradient():
    """Every interior face between two fluid cells changes by (p_left - p_right)/cp:
    the sweeps only ever add a discrete pressure gradient."""
    rng = np.random.default_rng(1)
    for prec in (64, 32):
        f = OracleFluid(3.0, 12, 9, 0.1, prec=prec)
        S = (rng.random((14, 11)) > 0.15).astype(np.float32)
        S[0, :] = S[-1, :] = S[:, 0] = S[:, -1] = 0
        f.set("S", S)
        U0 = rng.standard_normal((14, 11))
        V0 = rng.standard_normal((14, 11))
        f.set("U", U0)
        f.set("V", V0)
        f.run_pass("makeIncompressible", 0.05, 7)
        U, V, p = f.get("U"), f.get("V"), f.get("p")
        U0 = U0.astype(f.dtype)
        V0 = V0.astype(f.dtype)
        cp = 3.0 * 0.1 / 0.05
        fl = S != 0
        tol = 1e-9 if prec == 64 else 2e-4
        for i in range(1, 13):
            for j in range(1, 10):
                if fl[i, j] and fl[i - 1, j]:
                    assert abs((U[i, j] - U0[i, j]) - (p[i - 1, j] - p[i, j]) / cp) < tol
                if fl[i, j] and fl[i, j - 1]:
                    assert abs((V[i, j] - V0[i, j]) - (p[i, j - 1] - p[i, j]) / cp) < tol


def test_pressure_projection_reduces_divergence():
    # TestPressureProjection (fluid_test.go:210-264)
    f = fluid(8, 8)
    c = 5
    U = np.zeros((10, 10), np.float32)
    V = np.zeros((10, 10), np.float32)
    for i in range(1, 9):
        for j in range(1, 9):
            dx, dy = F32(i - c), F32(j - c)
            dist = F32(math.sqrt(float(dx * dx + dy * dy))) + F32(0.1)
            U[i, j], V[i, j] = dx / dist, dy / dist
    f.set("U", U)
    f.set("V", V)

    def avgdiv():
        U, V, S = f.get("U"), f.get("V"), f.get("S")
        d = U[2:10, 1:9] - U[1:9, 1:9] + V[1:9, 2:10] - V[1:9, 1:9]
        return float(np.abs(d[S[1:9, 1:9] > 0]).mean())
    d0 = avgdiv()
    f.run_pass("makeIncompressible", 0.01, 20)
    assert avgdiv() < 0.5 * d0
    assert f.last_sweeps() == 20


def test_early_exit_on_quiescent_field():
    f = fluid(16, 16)
    f.PressureIters = 40
    f.Simulate(0.1)
    assert f.last_sweeps() == 1
    st = f.state()
    assert not st["U"].any() and not st["V"].any() and not st["M"].any()


# ---- advection closed forms ----------------------------------------------------------------

def test_uniform_field_advects_to_itself():
    f = fluid(12, 10)
    f.set("U", np.full((14, 12), 0.7))
    f.set("V", np.full((14, 12), -0.3))
    f.set("M", np.full((14, 12), 0.4))
    f.run_pass("advectVelocity", 0.5)
    f.run_pass("advectSmoke", 0.5)
    assert np.allclose(f.get("U")[2:12, 2:10], 0.7, atol=1e-6)
    assert np.allclose(f.get("M")[2:12, 2:10], 0.4, atol=1e-6)


def test_integer_shift_copies_neighbour():
    """dt*u = h exactly: the smoke sample lands on cell centres (tx = 0)."""
    f = fluid(12, 10)
    U = np.ones((14, 12), np.float32)
    f.set("U", U)
    rng = np.random.default_rng(2)
    M = rng.random((14, 12)).astype(np.float32)
    f.set("M", M)
    f.run_pass("advectSmoke", 1.0)
    out = f.get("M")
    assert np.array_equal(out[2:12, 1:11], M[1:11, 1:11])


def test_bfecc_constant_field_identity():
    f = fluid(10, 8)
    f.set("U", np.full((12, 10), 0.25))
    f.set("M", np.full((12, 10), 0.6))
    f.run_pass("advectVelocityBFECC", 0.3)
    f.run_pass("advectSmokeBFECC", 0.3)
    assert np.allclose(f.get("U")[1:11, 1:9], 0.25, atol=1e-6)
    assert np.allclose(f.get("M")[1:11, 1:9], 0.6, atol=1e-6)


def test_velocity_and_smoke_advection_behaviour():
    # TestVelocityAdvection / TestSmokeAdvection (fluid_test.go:84-208)
    f = fluid(10, 10)
    U = np.zeros((12, 12), np.float32)
    U[1:11, 1:11] = 2.0
    V = np.zeros((12, 12), np.float32)
    V[3, 5] = 5.0
    f.set("U", U)
    f.set("V", V)
    f.run_pass("advectVelocity", 0.1)
    Vn = f.get("V")
    assert Vn[4, 5] > V[4, 5] and Vn[3, 5] < 4.5
    assert f.get("U")[1:11, 1:11].mean() >= 1.6
    g = fluid(10, 10)
    U = np.zeros((12, 12), np.float32)
    U[1:11, 1:11] = 3.0
    M = np.zeros((12, 12), np.float32)
    M[2, 5] = 1.0
    g.set("U", U)
    g.set("M", M)
    g.run_pass("advectSmoke", 0.1)
    Mn = g.get("M")
    assert any(Mn[i, 5] > M[i, 5] + 0.01 for i in range(2, 5))
    assert abs(float(Mn.sum()) - 1.0) <= 0.1


# ---- FP32 vs FP64 instantiations ------------------------------------------------------------

@pytest.mark.parametrize("bfecc,conf,tol", [(False, 0.0, 1e-5), (True, 0.0, 1e-5), (False, 0.3, 3e-4),
                                           (True, 0.3, 3e-4)])
def test_fp32_tracks_fp64(bfecc, conf, tol):
    """The two instantiations share control flow; their step results agree to float32
    rounding (a mis-typed constant or a missing term would not). Confinement normalises
    the |curl| gradient by |g| + 1e-5 (fluid.go:479-481), whose direction is
    ill-conditioned where |curl| is flat, so it amplifies rounding about tenfold."""
    rng = np.random.default_rng(3)
    outs = []
    gate = None
    for prec in (32, 64):
        f = OracleFluid(1000.0, 24, 20, 0.01, prec=prec)
        S = np.ones((26, 22), np.float32)
        S[0, :] = S[:, 0] = S[:, -1] = 0
        f.set("S", S)
        base = rng if prec == 32 else np.random.default_rng(3)
        f.set("U", (0.5 + 0.1 * base.standard_normal((26, 22))).astype(np.float32))
        f.set("V", (0.1 * base.standard_normal((26, 22))).astype(np.float32))
        f.set("M", base.random((26, 22)).astype(np.float32))
        f.UseBFECC = bfecc
        f.Confinement = conf
        f.PressureIters = 20
        if prec == 64:
            f.set_forced_sweeps(20)
            f.set_turb_gate(gate)
        f.Simulate(0.1)
        if prec == 32:
            gate = f.turb_gate()
        outs.append(f.state())
    for k in ("U", "V", "p", "M"):
        a, b = outs[0][k].astype(np.float64), outs[1][k]
        scale = max(np.abs(b).max(), 1e-30)
        assert np.abs(a - b).max() / scale < tol, k


# ---- application layer: colormaps and particles (main/colors.go, main.go:512-652) --------

def test_sci_colormap_closed_forms():
    # getSciValue (colors.go:46-84): min -> blue, the four quarter points, max clamped below red
    f = fluid(4, 1)  # 6 x 3 cells
    M = np.zeros((6, 3), F32)
    M[0, :] = [0.0, 0.25, 0.5]
    M[1, :] = [0.75, 1.0, 0.125]
    f.set("M", M)
    pix = f.render(0, overlay=False)        # image row j shows fluid row NumY-1-j
    px = lambda i, j: tuple(pix[f.NumY - 1 - j, i])  # noqa: E731
    assert px(0, 0) == (0, 0, 255, 255)          # val 0: num 0, s 0
    assert px(0, 1) == (0, 255, 255, 255)        # val 0.25: num 1, s 0 -> cyan
    assert px(0, 2) == (0, 255, 0, 255)          # val 0.5: num 2, s 0 -> green
    assert px(1, 0) == (255, 255, 0, 255)        # val 0.75: num 3, s 0 -> yellow
    # val 1 = max: clamped to 1 - 0.0001, num 3, s ~ 0.9996 -> (255, 0, 0) after truncation
    assert px(1, 1) == (255, int(F32(255) * (F32(1) - F32((F32(1) - F32(1e-4) - F32(0.75)) / F32(0.25)))), 0, 255)
    assert px(1, 2) == (0, 127, 255, 255)        # val 0.125: num 0, s 0.5 -> g = uint8(127.5)


def test_diverging_colormap_and_solid_overlay():
    f = fluid(6, 6, hh=1.0)
    U = np.zeros((8, 8), F32)
    U[3, 4] = 1.0   # curl at (3,3) and (3,5) has opposite signs
    f.set("U", U)
    S = f.get("S")
    S[6, 6] = 0
    f.set("S", S)
    vals = f.Vorticity()
    vmax = max(abs(vals.MinValue), abs(vals.MaxValue))
    pix = f.render(3, overlay=False)
    px = lambda i, j: tuple(pix[f.NumY - 1 - j, i])  # noqa: E731
    v = vals.values.reshape(8, 8)
    for (i, j) in [(3, 3), (3, 5), (2, 2)]:
        t = np.clip(F32(v[i, j]) / F32(vmax), -1, 1)
        if t >= 0:
            exp = (255, int(F32(255) * (F32(1) - t)), int(F32(255) * (F32(1) - t)), 255)
        else:
            a = -t
            exp = (int(F32(255) * (F32(1) - a)), int(F32(255) * (F32(1) - a)), 255, 255)
        assert px(i, j) == exp, (i, j, px(i, j), exp)
    assert px(2, 2) == (255, 255, 255, 255)      # zero curl -> white
    assert px(6, 6) == (0, 0, 0, 255)            # solid cell painted black


def test_particles_in_uniform_flow():
    # RK2 midpoint in a uniform field moves by exactly dt*u; age and survival rules
    f = fluid(10, 10, hh=1.0)
    f.set("U", np.full((12, 12), 0.5, F32))
    f.set("V", np.full((12, 12), -0.25, F32))
    x = np.array([3.5, 3.5, 10.9, 5.0], F32)
    y = np.array([4.5, 4.5, 5.0, 0.2], F32)
    age = np.array([0.0, 14.95, 0.0, 0.0], F32)
    mx = np.array([15.0, 15.0, 15.0, 15.0], F32)
    keep = f.advect_particles(F32(0.1), x, y, age, mx)
    assert x[0] == F32(3.5) + F32(0.1) * F32(0.5) and y[0] == F32(4.5) + F32(0.1) * F32(-0.25)
    assert keep.tolist() == [1, 0, 1, 1]            # particle 1 exceeded MaxAge
    assert age[1] == F32(14.95) + F32(0.1)
    S = f.get("S")
    nx, ny = x[2] + F32(0.1) * F32(0.5), y[2] + F32(0.1) * F32(-0.25)
    S[int(nx), int(ny)] = 0                        # the cell it reaches next is solid: dropped
    f.set("S", S)
    keep = f.advect_particles(F32(0.1), x, y, age, mx)
    assert keep[2] == 0 and x[2] == nx
