"""The reference's scenario tests (pkg/fluid/fluid_test.go), restated as functions of a
factory `make(density, width, height, h)`, so the same scenario drives the device (Fluid)
and the float32 oracle. Each returns (fluid, checks): `checks` maps the reference's
assertion to whether it held. Direct writes to the exported slices (f.U[k] = v), which the
reference's tests use, are done by read-modify-write of the whole field."""
import numpy as np

F32 = np.float32


def _write(f, name, cells, value):
    a = f.get(name)
    for (i, j) in cells:
        a[i, j] = value
    f.set(name, a)


def _all_fluid(f):
    f.set("S", np.ones((f.NumX, f.NumY), np.float32))


def _max_uv(f):
    U, V = np.abs(f.get("U")[1:-1, 1:-1]), np.abs(f.get("V")[1:-1, 1:-1])
    return float(U.max()), float(V.max())


def _final_divergence(f):
    """calculateFinalDivergence (fluid_test.go:1449-1468): mean |div| over fluid interior cells."""
    U, V, S = f.get("U"), f.get("V"), f.get("S")
    div = U[2:, 1:-1] - U[1:-1, 1:-1] + V[1:-1, 2:] - V[1:-1, 1:-1]
    m = S[1:-1, 1:-1] > 0
    return float(np.abs(div[m]).mean()) if m.any() else 0.0


def jet_rows(f, height):
    c = f.NumY // 2
    return [j for j in range(c - height // 2, c + height // 2) if 0 < j < f.NumY - 1]


def realistic_jet(make):
    """TestRealisticJetSimulation (fluid_test.go:357-477)."""
    f = make(1.0, 30, 20, 1.0)
    _all_fluid(f)
    rows, jx, v = jet_rows(f, 8), 1, 15.0
    _write(f, "U", [(jx, j) for j in rows], v)
    _write(f, "V", [(jx, j) for j in rows], 0.0)
    _write(f, "M", [(jx, j) for j in rows], 1.0)
    for _ in range(20):
        f.Simulate(0.05)
        _write(f, "U", [(jx, j) for j in rows], v)
        _write(f, "M", [(jx, j) for j in rows], 1.0)
    U, M = f.get("U"), f.get("M")
    checks = {"jet keeps > 0.6 of its speed": float(np.mean([U[jx + 1, j] for j in rows])) >= 0.6 * v}
    prev, spread_ok = 0.0, True
    for d in (3, 6, 10):
        x = jx + d
        if x >= f.NumX - 1:
            continue
        spread = float((U[x, 1:-1] > 0.1 * v).sum())
        if prev > 0 and spread < 0.9 * prev:
            spread_ok = False
        prev = spread
    checks["jet spreads downstream"] = spread_ok
    checks["smoke transported"] = bool((M[jx + 2:min(jx + 10, f.NumX - 1)][:, rows] > 0.1).any())
    checks["x momentum"] = float(U[1:-1, 1:-1].sum()) >= v * 8 * 1.0 * 0.3
    return f, checks


def obstacle_vortex_shedding(make):
    """TestObstacleVortexShedding (fluid_test.go:479-648)."""
    f = make(1.0, 40, 25, 1.0)
    _all_fluid(f)
    ox, oy, r = 15, f.NumY // 2, 3
    S = f.get("S")
    disc = [(i, j) for i in range(1, f.NumX - 1) for j in range(1, f.NumY - 1)
            if F32(i - ox) * F32(i - ox) + F32(j - oy) * F32(j - oy) <= F32(r * r)]
    for (i, j) in disc:
        S[i, j] = 0.0
    f.set("S", S)
    rows, jx, v = jet_rows(f, 6), 3, 12.0
    f.Confinement = 0.1
    _write(f, "M", [(jx, j) for j in rows], 1.0)
    for _ in range(120):
        f.Simulate(0.025)
        _write(f, "U", [(jx, j) for j in rows], v)
        _write(f, "M", [(jx, j) for j in rows], 1.0)
    U, V, M, p, S = (f.get(k) for k in ("U", "V", "M", "p", "S"))
    cr = r + 2
    box = [(i, j) for i in range(ox - cr, ox + cr + 1) for j in range(oy - cr, oy + cr + 1)
           if 0 < i < f.NumX - 1 and 0 < j < f.NumY - 1 and S[i, j] > 0]
    checks = {"obstacle deflects the flow": any(abs(V[i, j]) > 1.0 for (i, j) in box)}
    curl = 0.0
    for i in range(ox + r + 1, min(ox + 15, f.NumX - 2)):
        for j in range(2, f.NumY - 2):
            if S[i, j] > 0:
                c = (V[i + 1, j] - V[i - 1, j]) / F32(2.0) - (U[i, j + 1] - U[i, j - 1]) / F32(2.0)
                curl = max(curl, abs(float(c)))
    checks["wake vorticity"] = curl >= 0.2
    total, front = 0.0, -1
    for i in range(jx + 1, f.NumX - 1):
        has = False
        for j in range(1, f.NumY - 1):
            if S[i, j] > 0:
                total += float(M[i, j])
                if M[i, j] > 0.01:
                    has, front = True, i
        if not has and front > 0:
            break
    checks["smoke passes the obstacle"] = front >= ox - 2 and total >= 0.1
    checks["obstacle stays solid"] = all(S[i, j] == 0 for (i, j) in disc)
    near = [(i, j) for i in range(ox - 2, ox + 5) for j in range(oy - 2, oy + 3)
            if 0 < i < f.NumX - 1 and 0 < j < f.NumY - 1 and S[i, j] > 0]
    checks["pressure varies around the obstacle"] = any(abs(p[i, j]) > 0.1 for (i, j) in near)
    return f, checks


def force_keeps_incompressibility(make):
    """TestForceDoesNotBreakIncompressibility (fluid_test.go:1192-1208)."""
    f = make(1.0, 20, 15, 1.0)
    _all_fluid(f)
    f.ApplyForceRadius(10, 7, 20.0, 10.0, 4)
    f.Simulate(0.05)
    return f, {"max divergence <= 1": f.MaxDivergence() <= 1.0}


def bfecc_long_run(make):
    """TestBFECCStabilityLongRun (fluid_test.go:1270-1314)."""
    f = make(1.0, 30, 20, 1.0)
    _all_fluid(f)
    f.UseBFECC = True
    f.SetCircularObstacle(15, f.NumY // 2, 3)
    rows = [j for j in range(f.NumY // 2 - 4, f.NumY // 2 + 4) if 0 < j < f.NumY - 1]
    finite = True
    for _ in range(200):
        _write(f, "U", [(1, j) for j in rows], 15.0)
        _write(f, "M", [(1, j) for j in rows], 1.0)
        f.Simulate(0.05)
        for k in ("U", "V", "M"):
            finite = finite and bool(np.isfinite(f.get(k)[1:-1, 1:-1]).all())
    return f, {"finite for 200 steps": finite, "smoke >= -0.001": float(f.get("M").min()) >= -0.001}


def interaction_stability(make):
    """TestInteractionStability (fluid_test.go:1316-1358)."""
    f = make(1.0, 30, 20, 1.0)
    _all_fluid(f)
    ok = True
    for step in range(100):
        if step % 5 == 0:
            f.ApplyForceRadius(10 + step % 10, 10, 10.0, 5.0, 3)
        if step % 3 == 0:
            f.AddSmoke(5, 10, 1.0)
        if step % 20 == 0:
            f.SetSolid(15, 10, True)
        if step % 20 == 10:
            f.SetSolid(15, 10, False)
        f.Simulate(0.05)
        mu = max(_max_uv(f))
        if not (mu <= 500):
            ok = False
    return f, {"max velocity stays <= 500": ok}


def smoke_non_negative(make):
    """TestSmokeNonNegative (fluid_test.go:1360-1379)."""
    f = make(1.0, 20, 15, 1.0)
    _all_fluid(f)
    f.ViscosityDiffusion = 0.1
    for _ in range(50):
        _write(f, "U", [(3, 7)], 10.0)
        _write(f, "M", [(3, 7)], 1.0)
        f.Simulate(0.08)
    return f, {"smoke >= -0.01": float(f.get("M").min()) >= -0.01}


def visual_low_iterations(make):
    """TestVisualEnhancementsLowIterations (fluid_test.go:760-839)."""
    f = make(1.0, 20, 15, 1.0)
    _all_fluid(f)
    f.ViscosityDiffusion, f.PressureDamping, f.Confinement = 0.1, 0.95, 0.05
    rows, jx, v = jet_rows(f, 4), 2, 10.0
    _write(f, "S", [(10, f.NumY // 2)], 0.0)
    ok = True
    for _ in range(15):
        _write(f, "U", [(jx, j) for j in rows], v)
        _write(f, "M", [(jx, j) for j in rows], 1.0)
        f.Simulate(0.1)
        mu, mv = _max_uv(f)
        if not (mu <= 50 and mv <= 50):
            ok = False
    M = f.get("M")
    return f, {"velocities stay <= 50": ok, "smoke transported": float(M[jx + 2:-1, 1:-1].sum()) > 0}


def multigrid_stability(make):
    """TestMultigridStability (fluid_test.go:1470-1547)."""
    f = make(1.0, 30, 20, 1.0)
    _all_fluid(f)
    f.UseMultigrid, f.MultigridLevels = True, 2
    f.ViscosityDiffusion, f.PressureDamping, f.Confinement = 0.1, 0.95, 0.05
    c = f.NumY // 2
    rows, jx, v = jet_rows(f, 6), 3, 18.0
    _write(f, "S", [(15, c), (22, c - 2), (22, c + 2)], 0.0)
    ok = True
    for _ in range(25):
        _write(f, "U", [(jx, j) for j in rows], v)
        _write(f, "M", [(jx, j) for j in rows], 1.0)
        f.Simulate(0.08)
        mu, mv = _max_uv(f)
        if not (mu <= 100 and mv <= 100):
            ok = False
            break
    return f, {"velocities stay <= 100": ok, "final mean divergence <= 5": _final_divergence(f) <= 5.0}


def multigrid_performance(make, use_multigrid, iters):
    """TestMultigridPerformance (fluid_test.go:1024-1132): its customSimulate is Simulate with
    `iters` sweeps (V-cycles) and semi-Lagrangian advection."""
    f = make(1.0, 20, 15, 1.0)
    _all_fluid(f)
    f.UseMultigrid, f.MultigridLevels = use_multigrid, 2
    f.ViscosityDiffusion, f.PressureDamping, f.PressureIters = 0.1, 0.95, iters
    c = f.NumY // 2
    rows, jx, v = jet_rows(f, 4), 2, 15.0
    _write(f, "S", [(10, c)], 0.0)
    ok = True
    for _ in range(10):
        _write(f, "U", [(jx, j) for j in rows], v)
        _write(f, "M", [(jx, j) for j in rows], 1.0)
        f.Simulate(0.1)
        mu, mv = _max_uv(f)
        if not (mu <= 200 and mv <= 200):
            ok = False
            break
    return f, {"velocities stay <= 200": ok}


def diffusion_behavior(make):
    """TestDiffusionBehavior (fluid_test.go:37-82): one makeIncompressible(10, 0.01) pass on a
    spike spreads it to the four neighbours."""
    f = make(1.0, 10, 10, 1.0)
    _all_fluid(f)
    c = (f.NumX // 2, f.NumY // 2)
    _write(f, "U", [c], 10.0)
    _write(f, "V", [c], 10.0)
    f.run_pass("makeIncompressible", 0.01, 10)
    U, V = f.get("U"), f.get("V")
    checks = {"centre velocity reduced": float(U[c]) < 10.0}
    for (i, j) in ((c[0] - 1, c[1]), (c[0] + 1, c[1]), (c[0], c[1] - 1), (c[0], c[1] + 1)):
        if 0 < i < f.NumX - 1 and 0 < j < f.NumY - 1:
            checks["neighbour %d,%d gained velocity" % (i, j)] = abs(float(U[i, j])) >= 0.05 or abs(float(V[i, j])) >= 0.05
    return f, checks


def _stable(f, limit):
    mu, mv = _max_uv(f)
    return np.isfinite(mu) and np.isfinite(mv) and mu <= limit and mv <= limit


def bfecc_stability(make):
    """TestBFECCStability (fluid_test.go:650-702): uniform (2, 1) flow, dt 0.2, 100 steps, a kick
    at (5, 5) every 10 steps; the test leaves UseBFECC at its default."""
    f = make(1.0, 20, 20, 1.0)
    _all_fluid(f)
    inner = [(i, j) for i in range(1, f.NumX - 1) for j in range(1, f.NumY - 1)]
    _write(f, "U", inner, 2.0)
    _write(f, "V", inner, 1.0)
    ok = True
    for step in range(100):
        if not _stable(f, 1000.0):
            ok = False
            break
        f.Simulate(0.2)
        if step % 10 == 0:
            _write(f, "U", [(5, 5)], 3.0)
            _write(f, "V", [(5, 5)], 2.0)
    return f, {"velocities stay <= 1000 and finite": ok}


def bfecc_jet_stability(make):
    """TestBFECCJetStability (fluid_test.go:704-758): a 6-row jet at 20 held at i = 1 for 30
    steps of dt 0.1 (UseBFECC at its default)."""
    f = make(1.0, 30, 20, 1.0)
    _all_fluid(f)
    rows, jx = jet_rows(f, 6), 1
    ok = True
    for _ in range(30):
        _write(f, "U", [(jx, j) for j in rows], 20.0)
        _write(f, "V", [(jx, j) for j in rows], 0.0)
        if not _stable(f, 500.0):
            ok = False
            break
        f.Simulate(0.1)
    return f, {"velocities stay <= 500 and finite": ok}


def performance_comparison(make, viscosity, damping, confinement):
    """TestPerformanceComparison (fluid_test.go:841-899): a one-cell jet at 8 for 10 steps;
    max(|U| + |V|) over every index stays <= 100."""
    f = make(1.0, 15, 10, 1.0)
    _all_fluid(f)
    f.ViscosityDiffusion, f.PressureDamping, f.Confinement = viscosity, damping, confinement
    for _ in range(10):
        _write(f, "U", [(2, f.NumY // 2)], 8.0)
        f.Simulate(0.1)
    vel = np.abs(f.get("U")) + np.abs(f.get("V"))
    mx = float(vel.max())
    return f, {"max |U|+|V| <= 100 and finite": bool(np.isfinite(vel).all()) and mx <= 100.0}


def advanced_visual_enhancements(make):
    """TestAdvancedVisualEnhancements (fluid_test.go:901-1022): viscosity, damping, confinement,
    turbulence and SmokeAdvection 1.2 together, three solid cells, 20 steps at the adaptive
    time step from base 0.08."""
    f = make(1.0, 25, 15, 1.0)
    _all_fluid(f)
    f.ViscosityDiffusion, f.PressureDamping, f.Confinement = 0.15, 0.92, 0.08
    f.TurbulenceStrength, f.SmokeAdvection = 0.03, 1.2
    c = f.NumY // 2
    rows, jx = jet_rows(f, 6), 3
    _write(f, "S", [(12, c), (18, c - 3), (18, c + 3)], 0.0)
    ok = True
    for _ in range(20):
        dt = f.GetAdaptiveTimeStep(0.08)
        _write(f, "U", [(jx, j) for j in rows], 12.0)
        _write(f, "M", [(jx, j) for j in rows], 1.0)
        f.Simulate(dt)
        M = f.get("M")[1:-1, 1:-1]
        if not (_stable(f, 100.0) and np.isfinite(M).all()):
            ok = False
    downstream = float(f.get("M")[jx + 5:-1, 1:-1].astype(np.float64).sum())
    return f, {"velocities stay <= 100, no NaN": ok, "downstream smoke >= 2": downstream >= 2.0}


def bfecc_accuracy(make, use_bfecc):
    """TestBFECCAccuracy (fluid_test.go:1210-1268): a 5x5 smoke blob in a uniform flow of 3,
    50 steps of dt 0.05 with the flow restored after each. The reference only logs the peak
    smoke of the two methods; the check here is that it stays finite and within [0, 1]."""
    f = make(1.0, 30, 30, 1.0)
    _all_fluid(f)
    inner = [(i, j) for i in range(1, f.NumX - 1) for j in range(1, f.NumY - 1)]
    _write(f, "U", inner, 3.0)
    _write(f, "M", [(i, j) for i in range(8, 13) for j in range(13, 18)], 1.0)
    f.UseBFECC = use_bfecc
    for _ in range(50):
        f.Simulate(0.05)
        _write(f, "U", inner, 3.0)
    peak = float(f.get("M").max())
    return f, {"peak smoke finite, in [0, 1]": np.isfinite(peak) and 0.0 <= peak <= 1.0}


def go_benchmark(make, variant, frames=30):
    """BenchmarkSimulate, BenchmarkSimulateWithJet, BenchmarkSimulateBFECC
    (fluid_bench_test.go:19-60) run for `frames` iterations. The benchmarks assert nothing; the
    check is that the fields stay finite."""
    from fluid_amd import configs as C
    f = C.go_bench_fluid(make, variant)
    for _ in range(frames):
        C.go_bench_frame(f, variant)
    ok = all(np.isfinite(f.get(k)).all() for k in ("U", "V", "p", "M"))
    return f, {"fields finite": ok}


SCENARIOS = {
    "realistic_jet": realistic_jet,
    "obstacle_vortex_shedding": obstacle_vortex_shedding,
    "force_keeps_incompressibility": force_keeps_incompressibility,
    "bfecc_long_run": bfecc_long_run,
    "interaction_stability": interaction_stability,
    "smoke_non_negative": smoke_non_negative,
    "visual_low_iterations": visual_low_iterations,
    "multigrid_stability": multigrid_stability,
    "multigrid_performance_sg8": lambda mk: multigrid_performance(mk, False, 8),
    "multigrid_performance_mg4": lambda mk: multigrid_performance(mk, True, 4),
    "multigrid_performance_sg4": lambda mk: multigrid_performance(mk, False, 4),
    "diffusion_behavior": diffusion_behavior,
    "bfecc_stability": bfecc_stability,
    "bfecc_jet_stability": bfecc_jet_stability,
    "performance_comparison_enhanced": lambda mk: performance_comparison(mk, 0.1, 0.95, 0.05),
    "performance_comparison_standard": lambda mk: performance_comparison(mk, 0.0, 1.0, 0.0),
    "advanced_visual_enhancements": advanced_visual_enhancements,
    "bfecc_accuracy_sl": lambda mk: bfecc_accuracy(mk, False),
    "bfecc_accuracy_bfecc": lambda mk: bfecc_accuracy(mk, True),
    "go_benchmark_quiescent": lambda mk: go_benchmark(mk, "quiescent"),
    "go_benchmark_jet": lambda mk: go_benchmark(mk, "jet"),
    "go_benchmark_bfecc": lambda mk: go_benchmark(mk, "bfecc"),
}
