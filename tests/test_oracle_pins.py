"""One named pin per oracle function that the reference's own tests leave open.

Each pin checks the CPU oracle (oracle/fluid_oracle.cpp) against something the mathematics
fixes, not against a retyped copy of its formula: bilinear interpolation reproduces affine
fields; BFECC of an affine velocity is a composition of affine maps; the half-cell shift
reduces advection to neighbour averages, so the 3x3 clamp is checked by brute force; hand-
derived values for vorticity confinement on a 8x8 grid; the relaxation schedule recovered
from the divergence decay of one cell; exact spikes for viscosity and smoke diffusion;
the speed gate at chosen speeds; the multigrid pieces by constant, affine and quadratic
fields, a linear solve and a float32 loop; the CFL clamps; textbook Bresenham; presets and
sources from the scene description.

`test_planted_mistake_is_caught` then loads the plantable build of the same source and, for
each planted mistake (the list heads fluid_oracle.cpp), shows that its pin fails.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle
from oracle.pyoracle import OracleFluid

F32 = np.float32
EPS5 = float(F32(1e-5))
TENTH = float(F32(0.1))

PINS = {}


def pin(*precs):
    def deco(fn):
        PINS[fn.__name__] = (fn, precs)
        return fn
    return deco


def all_fluid(w, h, hh=1.0, prec=64, density=1.0):
    f = OracleFluid(density, w, h, hh, prec=prec)
    f.set("S", np.ones((f.NumX, f.NumY)))
    return f


def close(a, b, tol, what):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    err = np.max(np.abs(a - b)) if a.size else 0.0
    assert err <= tol, "%s: max error %.3g > %.3g" % (what, err, tol)


# ---- sampleField (fluid.go:357-398) ------------------------------------------------------

def affine_sample_expect(f, kind, x, y, a, b, c):
    """Clamped bilinear sample of data[i, j] = a*i + b*j + c at (x, y): x, y clamp to
    [h, N*h] (fluid.go:363-364); the node coordinate is position/h minus the kind's half-cell
    offset (U: y, V: x, M: both) and stops at N-1, where x1 = min(x0+1, N-1) repeats the node."""
    h = f.hval
    ox = 0.5 if kind in ("V", "M") else 0.0
    oy = 0.5 if kind in ("U", "M") else 0.0
    xc = min(max(x, h), f.NumX * h)
    yc = min(max(y, h), f.NumY * h)
    ex = min(xc / h - ox, f.NumX - 1)
    ey = min(yc / h - oy, f.NumY - 1)
    return a * ex + b * ey + c


@pin(64)
def pin_sample_affine(prec):
    """Bilinear sampling reproduces an affine field at any point, with the U/V/M half-cell
    offsets and both clamp bounds. A swapped corner pair gives a*ty + b*tx instead."""
    f = all_fluid(9, 7, hh=0.25, prec=prec)
    a, b, c = 1.75, -0.625, 3.0
    ii, jj = np.meshgrid(np.arange(f.NumX), np.arange(f.NumY), indexing="ij")
    data = a * ii + b * jj + c
    for name in ("U", "V", "M"):
        f.set(name, data)
    rng = np.random.default_rng(11)
    h = f.hval
    pts = list(zip(rng.uniform(-0.3, (f.NumX + 1) * h, 300), rng.uniform(-0.3, (f.NumY + 1) * h, 300)))
    pts += [(h, h), (f.NumX * h, f.NumY * h), (h, f.NumY * h), (f.NumX * h, h), (3 * h, 2.5 * h)]
    for kind in ("U", "V", "M"):
        for x, y in pts:
            x, y = float(F32(x)), float(F32(y))
            got = f.sample_field(kind, kind, x, y)
            want = affine_sample_expect(f, kind, x, y, a, b, c)
            assert abs(got - want) <= 1e-12, (kind, x, y, got, want)


@pin(32)
def pin_sample_affine_exact_float32(prec):
    """In float32, positions on an eighth-cell lattice and small integer coefficients keep
    every operation exact, so the sample equals the affine value bit for bit."""
    f = all_fluid(9, 7, hh=0.5, prec=prec)
    a, b, c = 3.0, -2.0, 5.0
    ii, jj = np.meshgrid(np.arange(f.NumX), np.arange(f.NumY), indexing="ij")
    for name in ("U", "V", "M"):
        f.set(name, a * ii + b * jj + c)
    h = f.hval
    for kind in ("U", "V", "M"):
        for x8 in range(0, 8 * f.NumX + 12, 3):
            for y8 in range(0, 8 * f.NumY + 12, 5):
                x, y = x8 * h / 8, y8 * h / 8
                got = f.sample_field(kind, kind, x, y)
                assert got == affine_sample_expect(f, kind, x, y, a, b, c), (kind, x, y)


# ---- BFECC (fluid.go:911-1051, 1094-1120) ------------------------------------------------

def face_positions(f):
    h = f.hval
    ii, jj = np.meshgrid(np.arange(f.NumX, dtype=np.float64), np.arange(f.NumY, dtype=np.float64), indexing="ij")
    return (ii * h, (jj + 0.5) * h), ((ii + 0.5) * h, jj * h), ((ii + 0.5) * h, (jj + 0.5) * h)


A_VEL = np.array([[0.05, -0.03], [0.02, -0.04]])
B_VEL = np.array([0.3, 0.2])


def vel(x, y):
    return A_VEL[0, 0] * x + A_VEL[0, 1] * y + B_VEL[0], A_VEL[1, 0] * x + A_VEL[1, 1] * y + B_VEL[1]


def bfecc_velocity_map(dt):
    """The scheme as maps of continuous space: forward F(x) = vel(x - dt vel(x)); backward
    with the original velocity bwd(x) = F(x + dt vel(x)); compensation C = 1.5 vel - 0.5 bwd;
    final forward pass traced by C itself: R(x) = C(x - dt C(x)). Every map of an affine field
    stays affine, and bilinear sampling and the face averages are exact on affine fields."""
    def F(x, y):
        u, v = vel(x, y)
        return vel(x - dt * u, y - dt * v)

    def bwd(x, y):
        u, v = vel(x, y)
        return F(x + dt * u, y + dt * v)

    def C(x, y):
        u, v = vel(x, y)
        bu, bv = bwd(x, y)
        return 1.5 * u - 0.5 * bu, 1.5 * v - 0.5 * bv

    def R(x, y):
        cu, cv = C(x, y)
        return C(x - dt * cu, y - dt * cv)
    return R, C


@pin(64)
def pin_bfecc_velocity_affine(prec):
    """advectVelocityBFECC of an affine velocity equals the composed maps of the scheme away
    from the border (the 3x3 clamp is inactive there: asserted). A backward trace with -dt or
    by the forward result's velocity, or a missing 1/2, changes the result."""
    f = all_fluid(24, 22, hh=0.5, prec=prec)
    (xu, yu), (xv, yv), _ = face_positions(f)
    f.set("U", vel(xu, yu)[0])
    f.set("V", vel(xv, yv)[1])
    dt = 0.375
    R, C = bfecc_velocity_map(dt)
    # precondition: the corrected value lies inside the 3x3 range of the original
    cu = C(xu, yu)[0]
    u0 = vel(xu, yu)[0]
    sl = (slice(1, -1), slice(1, -1))
    lo = np.min([np.roll(np.roll(u0, a, 0), b, 1) for a in (-1, 0, 1) for b in (-1, 0, 1)], axis=0)
    hi = np.max([np.roll(np.roll(u0, a, 0), b, 1) for a in (-1, 0, 1) for b in (-1, 0, 1)], axis=0)
    assert ((cu[sl] > lo[sl] + 1e-3) & (cu[sl] < hi[sl] - 1e-3)).all()
    f.run_pass("advectVelocityBFECC", dt)
    m = 6
    inner = (slice(m, f.NumX - m), slice(m, f.NumY - m))
    close(f.get("U")[inner], R(xu, yu)[0][inner], 1e-10, "BFECC U")
    close(f.get("V")[inner], R(xv, yv)[1][inner], 1e-10, "BFECC V")


@pin(64)
def pin_bfecc_smoke_affine(prec):
    """advectSmokeBFECC of an affine smoke field under an affine velocity: forward M(x - dt w),
    backward by the same (current) velocity with +dt, compensation 1.5 M - 0.5 bwd, final
    forward pass, with w the cell-centre velocity (exact average of two affine faces)."""
    f = all_fluid(24, 22, hh=0.5, prec=prec)
    (xu, yu), (xv, yv), (xm, ym) = face_positions(f)
    f.set("U", vel(xu, yu)[0])
    f.set("V", vel(xv, yv)[1])
    mfun = lambda x, y: 2.0 + 0.07 * x - 0.05 * y
    f.set("M", mfun(xm, ym))
    dt = 0.375

    def trace(g, s):
        def out(x, y):
            u, v = vel(x, y)
            return g(x + s * u, y + s * v)
        return out
    fwd = trace(mfun, -dt)
    bwd = trace(fwd, dt)
    corr = lambda x, y: 1.5 * mfun(x, y) - 0.5 * bwd(x, y)
    final = trace(corr, -dt)
    f.run_pass("advectSmokeBFECC", dt)
    m = 6
    inner = (slice(m, f.NumX - m), slice(m, f.NumY - m))
    close(f.get("M")[inner], final(xm, ym)[inner], 1e-10, "BFECC M")


def interp_clamped(row, pos, lo, hi):
    return np.interp(np.clip(pos, lo, hi), np.arange(row.size, dtype=np.float64), row)


def minmax3(a):
    """3x3 minimum and maximum over in-grid neighbours (brute force)."""
    nx, ny = a.shape
    lo, hi = a.copy(), a.copy()
    for i in range(nx):
        for j in range(ny):
            w = a[max(i - 1, 0):i + 2, max(j - 1, 0):j + 2]
            lo[i, j], hi[i, j] = w.min(), w.max()
    return lo, hi


@pin(32, 64)
def pin_bfecc_velocity_clamp(prec):
    """A velocity step u = 1/4 | 3/4 along x (v = 0), h = 1, dt = 2: every trace lands on a half
    or whole cell, so each pass is an exact 1-D interpolation along the rows and every value is
    dyadic. The compensated value overshoots the step and the 3x3 clamp must cut it."""
    f = all_fluid(20, 6, hh=1.0, prec=prec)
    nx, ny = f.NumX, f.NumY
    U = np.where(np.arange(nx) <= 10, 0.25, 0.75)[:, None] * np.ones((1, ny))
    f.set("U", U)
    dt = 2.0
    orig = U.copy()
    fwd = orig.copy()
    for j in range(1, ny - 1):
        for i in range(1, nx):
            fwd[i, j] = interp_clamped(orig[:, j], i - dt * orig[i, j], 1, nx - 1)
    bwd = np.zeros_like(orig)
    bwd[0, :], bwd[-1, :], bwd[:, 0], bwd[:, -1] = fwd[0, :], fwd[-1, :], fwd[:, 0], fwd[:, -1]
    for j in range(1, ny - 1):
        for i in range(1, nx):
            bwd[i, j] = interp_clamped(fwd[:, j], i + dt * orig[i, j], 1, nx - 1)
    corr = orig.copy()
    corr[1:-1, 1:-1] = orig[1:-1, 1:-1] - (bwd[1:-1, 1:-1] - orig[1:-1, 1:-1]) * 0.5
    unclamped = corr.copy()
    lo, hi = minmax3(orig)
    corr[1:-1, 1:-1] = np.clip(corr[1:-1, 1:-1], lo[1:-1, 1:-1], hi[1:-1, 1:-1])
    assert (corr != unclamped).any()  # the clamp is exercised
    want = corr.copy()
    for j in range(1, ny - 1):
        for i in range(1, nx):
            want[i, j] = interp_clamped(corr[:, j], i - dt * corr[i, j], 1, nx - 1)
    f.run_pass("advectVelocityBFECC", dt)
    assert np.array_equal(f.get("U").astype(np.float64), want)
    assert not f.get("V").any()


@pin(32, 64)
def pin_bfecc_smoke_clamp(prec):
    """A smoke spike moved half a cell per pass (u = 1/4, dt = 2, h = 1): forward and backward
    passes are exact neighbour averages; the compensation overshoots (1.25 at the spike, -1/8
    beside it) and the 3x3 clamp and max(., 0) bring it back. Brute force on the whole grid."""
    f = all_fluid(14, 10, hh=1.0, prec=prec)
    nx, ny = f.NumX, f.NumY
    f.set("U", np.full((nx, ny), 0.25))
    M = np.zeros((nx, ny))
    M[6, 5] = 1.0
    M[9, 3] = 0.5
    f.set("M", M)
    dt = 2.0

    def sl(src, s):
        out = src.copy()
        for i in range(1, nx - 1):
            for j in range(1, ny - 1):
                # cell centre i+1/2 traced by s*u, node coordinate = position - 1/2, clamped to
                # [1/2, N-1] (x >= h with the half-cell offset)
                out[i, j] = interp_clamped(src[:, j], i + s * 0.25, 0.5, nx - 1)
        return out
    fwd = np.maximum(sl(M, -dt), 0)
    bwd = np.zeros_like(M)
    bwd[0, :], bwd[-1, :], bwd[:, 0], bwd[:, -1] = fwd[0, :], fwd[-1, :], fwd[:, 0], fwd[:, -1]
    bwd[1:-1, 1:-1] = sl(fwd, dt)[1:-1, 1:-1]
    corr = M.copy()
    corr[1:-1, 1:-1] = M[1:-1, 1:-1] - (bwd[1:-1, 1:-1] - M[1:-1, 1:-1]) * 0.5
    unclamped = corr.copy()
    lo, hi = minmax3(M)
    corr[1:-1, 1:-1] = np.maximum(np.clip(corr[1:-1, 1:-1], lo[1:-1, 1:-1], hi[1:-1, 1:-1]), 0)
    assert unclamped[6, 5] == 1.25 and unclamped[5, 5] == -0.125 and corr[6, 5] == 1.0
    want = np.maximum(sl(corr, -dt), 0)
    f.run_pass("advectSmokeBFECC", dt)
    assert np.array_equal(f.get("M").astype(np.float64), want)


# ---- vorticity confinement (fluid.go:449-493) -------------------------------------------

@pin(32, 64)
def pin_confinement_hand_computed(prec):
    """8x8 grid, h = 1, all fluid. V = 2 on faces (4,3) and (5,3), U = 1.5 on face (3,3).
    Hand-computed curl (0.5/h central differences): 1 at (3,3) and (4,3), -1 at (5,3) and
    (6,3), 0.75 at (3,4), -0.75 at (3,2), 0 elsewhere. The normalised gradient of |curl|,
    N = g/(|g| + 1e-5), is non-zero where curl is: (3,3) N = (k, 0); (6,3) N = (-k, 0);
    (3,4) N = (0, -k); (3,2) N = (0, k), with k = 0.5/(0.5 + 1e-5); (4,3) and (5,3) have
    g = 0. With s = C (1 + 0.1 |(U,V)|) and U += s Ny curl dt, V -= s Nx curl dt, C = 2,
    dt = 1/4: V(3,3) = -2.3 k/4 (its speed is 1.5), V(6,3) = -2 k/4, U(3,4) = U(3,2) =
    -2*0.75 k/4. The flanks of the shear layer turn down: the force strengthens both
    counter-rotating vortices."""
    f = all_fluid(6, 6, hh=1.0, prec=prec)
    U = np.zeros((8, 8))
    V = np.zeros((8, 8))
    V[4, 3] = V[5, 3] = 2.0
    U[3, 3] = 1.5
    f.set("U", U)
    f.set("V", V)
    f.Confinement = 2.0
    f.run_pass("applyVorticityConfinement", 0.25)
    k = 0.5 / (0.5 + EPS5)
    Uw, Vw = U.copy(), V.copy()
    Vw[3, 3] = -2.0 * (1 + TENTH * 1.5) * k * 0.25
    Vw[6, 3] = -2.0 * k * 0.25
    Uw[3, 4] = Uw[3, 2] = -2.0 * 0.75 * k * 0.25
    tol = 1e-14 if prec == 64 else 3e-7
    close(f.get("U"), Uw, tol, "confinement U")
    close(f.get("V"), Vw, tol, "confinement V")


@pin(64)
def pin_confinement_spins_up_a_vortex(prec):
    """A counter-clockwise vortex u_theta = 0.2 r exp(-r^2/18): its vorticity
    0.2 exp(-r^2/18)(2 - r^2/9) is positive with |curl| falling off for r < 4.2 (a weaker
    negative ring lies outside). In that core N points inwards and N x curl is along the
    rotation: confinement spins the core up, so the tangential work over r < 3.5 is positive."""
    f = all_fluid(24, 24, hh=1.0, prec=prec)
    (xu, yu), (xv, yv), _ = face_positions(f)
    c = 13.0
    g = lambda x, y: np.exp(-((x - c) ** 2 + (y - c) ** 2) / 18.0)
    U0 = -(yu - c) * g(xu, yu) * 0.2
    V0 = (xv - c) * g(xv, yv) * 0.2
    f.set("U", U0)
    f.set("V", V0)
    f.Confinement = 0.5
    f.run_pass("applyVorticityConfinement", 0.1)
    dU, dV = f.get("U") - U0, f.get("V") - V0
    core_u = np.hypot(xu - c, yu - c) < 3.5
    core_v = np.hypot(xv - c, yv - c) < 3.5
    tang = (dU * -(yu - c))[core_u].sum() + (dV * (xv - c))[core_v].sum()
    assert tang > 0 and np.abs(dU[core_u]).max() > 1e-4


# ---- relaxation schedule (fluid.go:157-186) -----------------------------------------------

@pin(64)
def pin_omega_schedule(prec):
    """One interior cell with four fluid neighbours (s = 4) and divergence 1: each sweep maps
    div -> div (1 - omega_k) and adds cp * (-div/4) omega_k to p. The per-sweep max|div|
    therefore gives omega_k = 1 + |div_k+1|/|div_k|, which must be 1.9 - 0.7 k/N with the
    float32 constants 1.9 and 1.2 (the 'relaxation *= 0.8' line is dead)."""
    for N in (7, 20):
        f = all_fluid(1, 1, hh=1.0, prec=prec)
        U = np.zeros((3, 3))
        U[2, 1] = 1.0
        f.set("U", U)
        f.EarlyExit = False
        f.run_pass("makeIncompressible", 1.0, N)
        md = f.sweep_maxdiv()
        assert md.size == N
        hi, lo = float(F32(1.9)), float(F32(1.2))
        omega = np.array([hi - (hi - lo) * (k / N) for k in range(N)])
        close(1 + md[1:] / md[:-1], omega[:-1], 1e-9, "omega from the divergence decay")
        div = np.cumprod(np.concatenate([[1.0], 1 - omega[:-1]]))
        close(f.get("p")[1, 1], np.sum(-div / 4 * omega), 1e-12, "p sum")


# ---- viscosity and smoke diffusion (fluid.go:112-142, 421-426) -------------------------

@pin(32, 64)
def pin_viscosity_spike(prec):
    """A spike of 1 in U and -2 in V, nu = 1/8, dt = 1/2: each fluid neighbour gains
    nu*dt*spike, the spike loses 4 nu dt spike; a solid neighbour is left alone."""
    f = all_fluid(8, 8, hh=1.0, prec=prec)
    S = f.get("S")
    S[5, 5] = 0
    f.set("S", S)
    U = np.zeros((10, 10))
    V = np.zeros((10, 10))
    U[4, 5], V[4, 5] = 1.0, -2.0
    f.set("U", U)
    f.set("V", V)
    f.ViscosityDiffusion = 0.125
    f.run_pass("applyViscosity", 0.5)
    Uw, Vw = U.copy(), V.copy()
    for (i, j) in ((3, 5), (4, 4), (4, 6)):
        Uw[i, j], Vw[i, j] = 0.0625, -0.125
    Uw[4, 5], Vw[4, 5] = 0.75, -1.5
    assert np.array_equal(f.get("U"), Uw) and np.array_equal(f.get("V"), Vw)
    assert np.array_equal(f.get("newU"), Uw) and np.array_equal(f.get("newV"), Vw)


@pin(64)
def pin_smoke_diffusion_spike(prec):
    """At rest the smoke sample returns the cell itself; diffusion then adds
    nu * 0.3 * dt * (5-point Laplacian), with the float32 constant 0.3."""
    f = all_fluid(8, 8, hh=1.0, prec=prec)
    M = np.zeros((10, 10))
    M[4, 5] = 1.0
    f.set("M", M)
    f.ViscosityDiffusion = 0.125
    f.run_pass("advectSmoke", 0.5)
    d = 0.125 * float(F32(0.3)) * 0.5
    Mw = M.copy()
    for (i, j) in ((3, 5), (5, 5), (4, 4), (4, 6)):
        Mw[i, j] = d
    Mw[4, 5] = 1 - 4 * d
    close(f.get("M"), Mw, 1e-15, "smoke diffusion")


# ---- turbulence (fluid.go:496-526) -----------------------------------------------------------

@pin(64)
def pin_turbulence_gate(prec):
    """Noise is applied only where the speed sqrt(U^2+V^2) exceeds 0.1, scaled by
    min(speed/2, 1): speeds 0.05 (off), 0.4 (x0.2), 3 (x1), (0.06, 0.05) -> 0.078 (off, though
    |U|+|V| = 0.11), (0.09, -0.09) -> 0.127 (on); a solid cell is left alone."""
    f = all_fluid(10, 10, hh=1.0, prec=prec)
    S = f.get("S")
    S[3, 3] = 0
    f.set("S", S)
    cells = {(2, 3): (0.05, 0.0), (4, 4): (0.24, 0.32), (6, 5): (3.0, 0.0), (7, 7): (0.06, 0.05),
             (8, 2): (0.09, -0.09), (3, 3): (1.0, 1.0)}
    U = np.zeros((12, 12))
    V = np.zeros((12, 12))
    for (i, j), (u, v) in cells.items():
        U[i, j], V[i, j] = u, v
    f.set("U", U)
    f.set("V", V)
    f.TurbulenceStrength = 0.02
    dt = 0.5
    f.run_pass("addTurbulence", dt)
    ts = float(F32(0.02)) * dt
    Uw, Vw = U.copy(), V.copy()
    for (i, j), (u, v) in cells.items():
        sp = math.hypot(u, v)
        if (i, j) == (3, 3) or sp <= 0.1:
            continue
        fac = min(0.5 * sp, 1.0)
        nu = float(F32(math.sin(float(F32(i * 137 + j * 241) * F32(0.01)))))
        nv = float(F32(math.sin(float(F32(i * 157 + j * 263) * F32(0.01)))))
        Uw[i, j] += nu * ts * fac
        Vw[i, j] += nv * ts * fac
    close(f.get("U"), Uw, 1e-15, "turbulence U")
    close(f.get("V"), Vw, 1e-15, "turbulence V")


# ---- GetAdaptiveTimeStep (fluid.go:529-557) ---------------------------------------------

@pin(64)
def pin_adaptive_time_step(prec):
    """dt = 0.8 h / max(|U|+|V|) over fluid interior cells, clamped to [0.1, 2] x basedt;
    basedt when at rest."""
    h, base = 0.01, float(F32(0.1))
    hf = float(F32(h))
    cases = [({}, base), ({(3, 3): (0.03, 0.02)}, 0.8 * hf / 0.05), ({(3, 3): (0.03, 0.04)}, 0.8 * hf / 0.07),
             ({(3, 3): (0.01, 0.01)}, 2 * base), ({(3, 3): (7.0, -3.0)}, float(F32(0.1)) * base),
             ({(3, 3): (0.03, 0.02), (0, 4): (50.0, 0.0), (5, 5): (40.0, 0.0)}, 0.8 * hf / 0.05)]
    for cells, want in cases:
        f = all_fluid(8, 8, hh=h, prec=prec)
        S = f.get("S")
        S[5, 5] = 0
        f.set("S", S)
        U = np.zeros((10, 10))
        V = np.zeros((10, 10))
        for (i, j), (u, v) in cells.items():
            U[i, j], V[i, j] = float(F32(u)), float(F32(v))
        f.set("U", U)
        f.set("V", V)
        mx = max([abs(U[i, j]) + abs(V[i, j]) for (i, j) in cells if (i, j) not in ((0, 4), (5, 5))] + [0])
        if mx:
            want = min(max(float(F32(0.8)) * hf / mx, float(F32(0.1)) * base), 2 * base)
        got = f.GetAdaptiveTimeStep(base)
        assert abs(got - want) <= 1e-12 * want, (cells, got, want)


# ---- multigrid (fluid.go:601-758, 1123-1149) ----------------------------------------------

@pin(64)
def pin_mg_residual_quadratic(prec):
    """r = -div - lap p: with U = a i, V = b j (div = a + b) and p = al i^2 + be j^2 on an
    all-fluid grid (lap p = 2 al + 2 be), r = -(a + b) - 2(al + be) on every interior cell; a
    solid cell has r = 0."""
    f = all_fluid(9, 8, hh=1.0, prec=prec)
    ii, jj = np.meshgrid(np.arange(f.NumX, dtype=float), np.arange(f.NumY, dtype=float), indexing="ij")
    a, b, al, be = 0.5, -1.25, 0.375, 0.125
    f.set("U", a * ii)
    f.set("V", b * jj)
    f.set("p", al * ii ** 2 + be * jj ** 2)
    r = f.mg_residual()
    want = np.zeros_like(r)
    want[1:-1, 1:-1] = -(a + b) - 2 * (al + be)
    assert np.array_equal(r, want)
    S = f.get("S")
    S[4, 4] = 0
    f.set("S", S)
    assert f.mg_residual()[4, 4] == 0


@pin(32, 64)
def pin_mg_restriction(prec):
    """Full weighting (1/4, 1/8, 1/16) sums to 1: a constant restricts to itself on the
    coarse interior; a linear field restricts to its value at the fine node (2i, 2j)."""
    f = all_fluid(13, 11, hh=1.0, prec=prec)
    cnx, cny = f.mg_dims()
    inner = np.zeros((cnx, cny), bool)
    for i in range(1, cnx - 1):
        for j in range(1, cny - 1):
            inner[i, j] = 2 * i < f.NumX - 1 and 2 * j < f.NumY - 1
    c = f.mg_restrict(np.full((f.NumX, f.NumY), 0.75))
    assert (c[inner] == 0.75).all() and not c[~inner].any()
    ii, jj = np.meshgrid(np.arange(f.NumX, dtype=float), np.arange(f.NumY, dtype=float), indexing="ij")
    c = f.mg_restrict(0.5 * ii - 0.25 * jj + 1)
    ci, cj = np.meshgrid(np.arange(cnx, dtype=float), np.arange(cny, dtype=float), indexing="ij")
    assert np.array_equal(c[inner], (0.5 * 2 * ci - 0.25 * 2 * cj + 1)[inner])


def coarse_gs_float32(rhs, cs):
    """solveCoarseGrid restated as plain scalar float32 loops (40 sweeps, relaxation 1.6)."""
    cnx, cny = rhs.shape
    P = np.zeros((cnx, cny), F32)
    rel = F32(1.6)
    for _ in range(40):
        for i in range(1, cnx - 1):
            for j in range(1, cny - 1):
                if cs[i, j] == 0:
                    continue
                sx0, sx1, sy0, sy1 = cs[i - 1, j], cs[i + 1, j], cs[i, j - 1], cs[i, j + 1]
                s = F32(F32(F32(sx0 + sx1) + sy0) + sy1)
                if s == 0:
                    continue
                pc = P[i, j]
                lap = F32(F32(F32(sx0 * F32(P[i - 1, j] - pc)) + F32(sx1 * F32(P[i + 1, j] - pc))) +
                          F32(sy0 * F32(P[i, j - 1] - pc)))
                lap = F32(lap + F32(sy1 * F32(P[i, j + 1] - pc)))
                res = F32(rhs[i, j] - lap)
                P[i, j] = F32(pc + F32(F32(-res / s) * rel))
    return P


@pin(32)
def pin_mg_coarse_solve_float32_loop(prec):
    """The coarse Gauss-Seidel equals a scalar float32 loop bit for bit on a 6x6 coarse grid
    with a solid coarse cell (coarse mask = fine S at even nodes)."""
    f = all_fluid(9, 9, hh=1.0, prec=prec)
    S = f.get("S")
    S[4, 6] = 0
    f.set("S", S)
    cnx, cny = f.mg_dims()
    assert (cnx, cny) == (6, 6)
    rhs = np.random.default_rng(5).standard_normal((cnx, cny)).astype(F32)
    cs = S[::2, ::2].astype(F32)
    assert np.array_equal(f.mg_coarse_solve(rhs), coarse_gs_float32(rhs, cs))


@pin(64)
def pin_mg_coarse_solve_converges(prec):
    """After its 40 sweeps the coarse solve satisfies lap P = rhs (P = 0 on the ring) to
    rounding: compared with a direct solve of the same 5-point system."""
    f = all_fluid(9, 9, hh=1.0, prec=prec)
    cnx, cny = f.mg_dims()
    rhs = np.random.default_rng(6).standard_normal((cnx, cny))
    n = (cnx - 2) * (cny - 2)
    A = np.zeros((n, n))
    idx = lambda i, j: (i - 1) * (cny - 2) + (j - 1)
    for i in range(1, cnx - 1):
        for j in range(1, cny - 1):
            A[idx(i, j), idx(i, j)] = -4
            for a, b in ((i - 1, j), (i + 1, j), (i, j - 1), (i, j + 1)):
                if 1 <= a < cnx - 1 and 1 <= b < cny - 1:
                    A[idx(i, j), idx(a, b)] = 1
    x = np.linalg.solve(A, rhs[1:-1, 1:-1].reshape(-1))
    close(f.mg_coarse_solve(rhs)[1:-1, 1:-1].reshape(-1), x, 1e-7 * np.abs(x).max(), "coarse solve")


@pin(32, 64)
def pin_mg_prolongation_affine(prec):
    """Bilinear prolongation of an affine coarse field is the same affine function at fine
    node (i/2, j/2), for fine nodes whose coarse cell has a right and an upper neighbour."""
    f = all_fluid(13, 10, hh=1.0, prec=prec)
    cnx, cny = f.mg_dims()
    ci, cj = np.meshgrid(np.arange(cnx, dtype=float), np.arange(cny, dtype=float), indexing="ij")
    a, b, c = 1.5, -0.75, 2.0
    fine = f.mg_prolong(a * ci + b * cj + c)
    ii, jj = np.meshgrid(np.arange(f.NumX, dtype=float), np.arange(f.NumY, dtype=float), indexing="ij")
    want = a * ii / 2 + b * jj / 2 + c
    ok = (ii // 2 < cnx - 1) & (jj // 2 < cny - 1)
    assert np.array_equal(fine[ok], want[ok]) and not fine[~ok].any()


@pin(32, 64)
def pin_mg_correction_gradient(prec):
    """applePressureCorrection adds cp*corr to p and the discrete gradient -d(corr) to every
    face between two fluid cells; faces against a solid wall do not move. Affine corr gives
    exact constants."""
    f = OracleFluid(1.0, 7, 6, 1.0, prec=prec)
    S = np.ones((f.NumX, f.NumY))
    S[0, :] = S[-1, :] = S[:, 0] = S[:, -1] = 0
    f.set("S", S)
    ii, jj = np.meshgrid(np.arange(f.NumX, dtype=float), np.arange(f.NumY, dtype=float), indexing="ij")
    a, b = 0.25, -0.5
    corr = a * ii + b * jj + 1
    f.mg_correct(corr, 2.0)
    U, V, p = f.get("U"), f.get("V"), f.get("p")
    Uw = np.zeros_like(U)
    Uw[2:-1, 1:-1] = -a
    Vw = np.zeros_like(V)
    Vw[1:-1, 2:-1] = -b
    pw = np.zeros_like(p)
    pw[1:-1, 1:-1] = 2.0 * corr[1:-1, 1:-1]
    assert np.array_equal(U, Uw) and np.array_equal(V, Vw) and np.array_equal(p, pw)


# ---- scene layer: walls, strokes, presets, sources, brush (main.go:453-486, 747-858) ------

def bresenham_cells(x0, y0, x1, y1):
    """Textbook Bresenham by major axis: one cell per major step, the minor coordinate rounded
    to nearest with exact ties moving towards the end point."""
    dx, dy = x1 - x0, y1 - y0
    L = max(abs(dx), abs(dy))
    if L == 0:
        return [(x0, y0)]
    sx, sy = (1 if dx > 0 else -1), (1 if dy > 0 else -1)
    cells = []
    for t in range(L + 1):
        if abs(dx) >= abs(dy):
            cells.append((x0 + sx * t, y0 + sy * ((2 * abs(dy) * t + L) // (2 * L))))
        else:
            cells.append((x0 + sx * ((2 * abs(dx) * t + L) // (2 * L)), y0 + sy * t))
    return cells


@pin(32)
def pin_draw_line_bresenham(prec):
    """drawLine's stroke equals the textbook cell list in all 8 octants, both directions, on
    axis-aligned and diagonal lines and ties; cells outside the grid are skipped."""
    rng = np.random.default_rng(7)
    segs = [(10, 10, 17, 13), (10, 10, 13, 17), (10, 10, 3, 13), (10, 10, 7, 17), (10, 10, 3, 7),
            (10, 10, 7, 3), (10, 10, 17, 7), (10, 10, 13, 3), (5, 5, 5, 15), (5, 5, 15, 5), (5, 5, 12, 12),
            (12, 5, 5, 12), (2, 2, 6, 4), (6, 4, 2, 2), (2, 2, 4, 6), (0, 0, 0, 0), (-5, 3, 8, 9)]
    segs += [tuple(int(v) for v in rng.integers(0, 22, 4)) for _ in range(60)]
    for x0, y0, x1, y1 in segs:
        f = OracleFluid(1.0, 20, 20, 1.0, prec=prec)
        f.set("S", np.ones((22, 22)))
        f.draw_line(x0, y0, x1, y1, True)
        got = {(int(i), int(j)) for i, j in zip(*np.nonzero(f.get("S") == 0))}
        want = {(i, j) for i, j in bresenham_cells(x0, y0, x1, y1) if 0 <= i < 22 and 0 <= j < 22}
        assert got == want, (x0, y0, x1, y1, sorted(got ^ want))


def expected_mask(nx, ny, top, bottom, left, right, disc=None):
    """All cells fluid, the preset's disc solid, then applyWallSettings: the bottom (j = 0) and
    top (j = NY-1) rows, then the left and right columns, which decide the corners."""
    S = np.ones((nx, ny))
    if disc:
        cx, cy, r = disc
        for i in range(nx):
            for j in range(ny):
                if (i - cx) ** 2 + (j - cy) ** 2 <= r * r:
                    S[i, j] = 0
    S[:, 0] = 0 if bottom else 1
    S[:, ny - 1] = 0 if top else 1
    S[0, :] = 0 if left else 1
    S[nx - 1, :] = 0 if right else 1
    return S


@pin(32)
def pin_presets_masks_and_sources(prec):
    """Jet: walls top, bottom, left; the open right column makes its two corners fluid.
    Cavity: all walls plus one source per cell of row NY-2, i in [2, NX-3], u = lid.
    Karman: jet walls plus a radius-8 disc at (NX/4, NY/2). Reset clears every field but S."""
    nx, ny = 42, 32
    f = OracleFluid(1.0, nx - 2, ny - 2, 1.0, prec=prec)
    f.set("M", np.ones((nx, ny)))
    f.reset_to_preset("jet")
    assert np.array_equal(f.get("S"), expected_mask(nx, ny, True, True, True, False))
    assert not f.get("M").any() and f.num_sources() == 0
    f.reset_to_preset("cavity", 2.5)
    assert np.array_equal(f.get("S"), expected_mask(nx, ny, True, True, True, True))
    assert f.num_sources() == nx - 4 and f.num_sources(sink=True) == 0
    f.apply_sources()
    U, V, M = f.get("U"), f.get("V"), f.get("M")
    Uw = np.zeros((nx, ny))
    Mw = np.zeros((nx, ny))
    Uw[2:nx - 2, ny - 2] = 2.5
    Mw[2:nx - 2, ny - 2] = 0.5
    assert np.array_equal(U, Uw) and not V.any() and np.array_equal(M, Mw)
    f.reset_to_preset("karman")
    assert np.array_equal(f.get("S"), expected_mask(nx, ny, True, True, True, False, disc=(nx // 4, ny // 2, 8)))


@pin(32)
def pin_sources_skip_solid_cells(prec):
    """applySources sets the velocity and adds 0.5 smoke at sources, only the velocity at sinks,
    and skips every source or sink on a solid cell."""
    nx, ny = 30, 20
    f = OracleFluid(1.0, nx - 2, ny - 2, 1.0, prec=prec)
    f.reset_to_preset("cavity", 3.0)
    f.draw_line(5, ny - 2, 9, ny - 2, True)
    f.add_sink(12, 5, -1.0, 0.5)
    f.add_sink(0, 5, 7.0, 7.0)
    f.apply_sources()
    U, V, M = f.get("U"), f.get("V"), f.get("M")
    for i in range(2, nx - 2):
        solid = 5 <= i <= 9
        assert U[i, ny - 2] == (0 if solid else 3.0) and M[i, ny - 2] == (0 if solid else 0.5), i
    assert U[12, 5] == -1.0 and V[12, 5] == 0.5 and M[12, 5] == 0
    assert U[0, 5] == 0 and V[0, 5] == 0


@pin(32)
def pin_smoke_brush(prec):
    """applySmokeBrush adds 0.3*strength on the disc dx^2 + dy^2 <= r^2 (29 cells at r = 3)."""
    f = OracleFluid(1.0, 18, 18, 1.0, prec=prec)
    f.smoke_brush(7, 9, 3, 2.0)
    M = f.get("M")
    want = np.zeros_like(M)
    for i in range(20):
        for j in range(20):
            if (i - 7) ** 2 + (j - 9) ** 2 <= 9:
                want[i, j] = F32(2.0) * F32(0.3)
    assert np.array_equal(M, want) and int((M != 0).sum()) == 29


# ---- the tests ------------------------------------------------------------------------------

CASES = [(n, p) for n, (fn, precs) in PINS.items() for p in precs]


@pytest.mark.parametrize("name,prec", CASES, ids=["%s-fp%d" % c for c in CASES])
def test_oracle_pin(name, prec):
    PINS[name][0](prec)


PLANTED = [
    (1, "pin_sample_affine"), (2, "pin_sample_affine"), (1, "pin_sample_affine_exact_float32"),
    (3, "pin_bfecc_velocity_affine"), (4, "pin_bfecc_velocity_affine"), (7, "pin_bfecc_velocity_affine"),
    (5, "pin_bfecc_smoke_affine"), (7, "pin_bfecc_smoke_affine"), (6, "pin_bfecc_velocity_clamp"),
    (6, "pin_bfecc_smoke_clamp"),
    (8, "pin_confinement_hand_computed"), (9, "pin_confinement_hand_computed"),
    (8, "pin_confinement_spins_up_a_vortex"), (10, "pin_omega_schedule"), (11, "pin_viscosity_spike"),
    (12, "pin_smoke_diffusion_spike"), (13, "pin_turbulence_gate"), (14, "pin_turbulence_gate"),
    (20, "pin_adaptive_time_step"), (21, "pin_adaptive_time_step"), (15, "pin_mg_residual_quadratic"),
    (16, "pin_mg_restriction"), (17, "pin_mg_coarse_solve_float32_loop"), (18, "pin_mg_prolongation_affine"),
    (19, "pin_mg_correction_gradient"), (22, "pin_draw_line_bresenham"), (23, "pin_presets_masks_and_sources"),
    (25, "pin_presets_masks_and_sources"), (26, "pin_presets_masks_and_sources"),
    (24, "pin_sources_skip_solid_cells"),
]


@pytest.mark.parametrize("plant,name", PLANTED, ids=["plant%d-%s" % c for c in PLANTED])
def test_planted_mistake_is_caught(plant, name):
    fn, precs = PINS[name]
    with pyoracle.planted(plant):
        for prec in precs:
            with pytest.raises(AssertionError):
                fn(prec)


def test_plantable_build_is_the_oracle_with_no_plant():
    """With no plant switched on, the plantable build passes every pin."""
    with pyoracle.planted(0):
        for name, (fn, precs) in PINS.items():
            for prec in precs:
                fn(prec)
