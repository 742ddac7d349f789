"""Synthetic workloads: the jet, Karman, cavity and differentiable-rollout scenes.

Each builder takes a factory `make(density, width, height, h)` returning any object
with the pkg/fluid API (fluid_amd.Fluid, or the oracle's mirror in tests) and sets up
the scene through that API only, so two implementations driven by the same builder
start from bit-identical states. Randomness comes from SplitMix64, a counter-based
generator that is easy to restate in any language (seed + (k+1)*0x9E3779B97F4A7C15,
then the two xor-shift-multiply rounds).

Scene constants not fixed by the workload description are recorded here:
density 1000 and h = 0.01 (the application's New(1000, ..., 1/100), main.go:201),
dt = 0.1, jet rows [H/2-8, H/2+8) at i = 1.
"""
import numpy as np

MASK = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


def splitmix64(seed, n):
    """First n outputs of SplitMix64 seeded with `seed`, as uint64."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & MASK) + k * np.uint64(GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform01(seed, n):
    """n float64 samples in [0,1) (top 53 bits)."""
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


class Rng:
    """Sequential SplitMix64 stream."""

    def __init__(self, seed):
        self.state = seed & MASK

    def u64(self):
        self.state = (self.state + GOLDEN) & MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self):
        return (self.u64() >> 11) * (1.0 / (1 << 53))

    def randint(self, lo, hi):
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1))


DENSITY = 1000.0
H = 0.01
DT = 0.1


def open_all_cells(f):
    """NewGame's initialisation: every cell fluid (main.go:222-226)."""
    S = np.ones((f.NumX, f.NumY), dtype=np.float32)
    f.set("S", S)


GO_BENCH_VARIANTS = ("quiescent", "jet", "bfecc")
GO_BENCH_DT = 1.0 / 120.0


def go_bench_fluid(make, variant="jet"):
    """newBenchFluid (fluid_bench_test.go:5-17): New(1000, 300, 251, 0.01), border cells solid,
    the interior fluid; Simulate's own 8 sweeps. The "bfecc" variant sets UseBFECC
    (BenchmarkSimulateBFECC, fluid_bench_test.go:45-60)."""
    f = make(1000.0, 300, 251, 0.01)
    S = np.zeros((f.NumX, f.NumY), dtype=np.float32)
    S[1:-1, 1:-1] = 1.0
    f.set("S", S)
    f.PressureIters = 8
    f.UseBFECC = variant == "bfecc"
    return f


def go_bench_frame(f, variant="jet"):
    """One benchmark iteration: BenchmarkSimulate (fluid_bench_test.go:19-27) is Simulate(1/120)
    alone; BenchmarkSimulateWithJet and BenchmarkSimulateBFECC (:29-60) first write U = 4 and
    M = 1 on the 100 cells i = 1, j in [NumY/2 - 50, NumY/2 + 50)."""
    if variant != "quiescent":
        lo = f.NumY // 2 - 50
        f.fill_cells("U", 1, 2, lo, lo + 100, 4.0)
        f.fill_cells("M", 1, 2, lo, lo + 100, 1.0)
    f.Simulate(GO_BENCH_DT)


def jet_scene(make, width=256, height=256, iters=40, jet_u=1.0, jet_rows=16, walls=(True, True, True, True)):
    """Jet: inlet at the left edge, zero initial state, all border walls."""
    f = make(DENSITY, width, height, H)
    open_all_cells(f)
    f.scene_walls(*walls)
    f.apply_wall_settings()
    lo = height // 2 - jet_rows // 2
    f.set_jet(True, 1, lo, lo + jet_rows, jet_u, 1.0)
    f.PressureIters = iters
    return f


def karman_scene(make, width=2048, height=1024, iters=100, radius=64, centre=(256, 512), seed=2, band=32):
    """Karman street: disc obstacle, uniform inflow u=1 on the left, open right edge, BFECC,
    confinement 0.3, seeded smoke noise in the inflow band i in [1, band]."""
    f = make(DENSITY, width, height, H)
    open_all_cells(f)
    f.scene_walls(True, True, True, False)
    f.apply_wall_settings()
    f.SetCircularObstacle(centre[0], centre[1], radius)
    for j in range(1, f.NumY - 1):
        f.add_sink(1, j, 1.0, 0.0)
    M = np.zeros((f.NumX, f.NumY), dtype=np.float32)
    noise = uniform01(seed, band * (f.NumY - 2)).astype(np.float32).reshape(band, f.NumY - 2)
    M[1:band + 1, 1:f.NumY - 1] = noise
    f.set("M", M)
    f.UseBFECC = True
    f.Confinement = 0.3
    f.PressureIters = iters
    return f


def cavity_discs(nx, ny, seed=3, frac=0.05, rmin=8, rmax=64):
    """Seeded discs (radius uniform in rmin..rmax, centre uniform in the interior) added
    until at least `frac` of the interior cells are solid. Returns [(cx, cy, r)]."""
    rng = Rng(seed)
    solid = np.zeros((nx, ny), dtype=bool)
    interior = (nx - 2) * (ny - 2)
    discs = []
    ii = np.arange(nx)[:, None]
    jj = np.arange(ny)[None, :]
    while solid[1:nx - 1, 1:ny - 1].sum() < frac * interior:
        r = rng.randint(rmin, rmax)
        cx = rng.randint(1, nx - 2)
        cy = rng.randint(1, ny - 2)
        x0, x1 = max(cx - r, 0), min(cx + r, nx - 1)
        y0, y1 = max(cy - r, 0), min(cy + r, ny - 1)
        dx = (ii[x0:x1 + 1] - cx).astype(np.float32)
        dy = (jj[:, y0:y1 + 1] - cy).astype(np.float32)
        solid[x0:x1 + 1, y0:y1 + 1] |= (dx * dx + dy * dy) <= np.float32(r * r)
        discs.append((cx, cy, r))
    return discs


def cavity_scene(make, width=4096, height=4096, iters=200, lid_u=1.0, seed=3, frac=0.05, bfecc=True):
    """Lid-driven cavity. All walls; the lid is the source row j = NY-2,
    i in [2, NX-3] (resetToPreset's cavity, main.go:767-779) with U = lid_u; seeded
    solid discs cover >= 5% of the interior; BFECC on."""
    f = make(DENSITY, width, height, H)
    open_all_cells(f)
    f.scene_walls(True, True, True, True)
    f.apply_wall_settings()
    for (cx, cy, r) in cavity_discs(f.NumX, f.NumY, seed, frac, 8, max(8, min(64, min(width, height) // 8))):
        f.SetCircularObstacle(cx, cy, r)
    for i in range(2, f.NumX - 2):
        f.add_source(i, f.NumY - 2, lid_u, 0.0)
    f.UseBFECC = bfecc
    f.PressureIters = iters
    return f


def diff_scene(make, width=512, height=512, iters=80):
    """Differentiable-rollout scene: the jet at 512^2 with BFECC and confinement 0.3 (callers
    of the rollout lower it: 0.3 diverges on this grid)."""
    f = jet_scene(make, width, height, iters)
    f.UseBFECC = True
    f.Confinement = 0.3
    return f


def gaussian_blobs(nx, ny, seed=5, count=6, sigma=20.0):
    """Target smoke field of the differentiable rollout: sum of `count` seeded Gaussians (sigma cells)."""
    rng = Rng(seed)
    ii = np.arange(nx, dtype=np.float64)[:, None]
    jj = np.arange(ny, dtype=np.float64)[None, :]
    T = np.zeros((nx, ny))
    for _ in range(count):
        cx = 1 + rng.uniform() * (nx - 2)
        cy = 1 + rng.uniform() * (ny - 2)
        T += np.exp(-((ii - cx) ** 2 + (jj - cy) ** 2) / (2 * sigma * sigma))
    return T.astype(np.float32)


# ---- the batched ensemble --------------------------------------------------------------
ENSEMBLE_SIDES = tuple(range(64, 1025, 32))  # {64, 96, ..., 1024}
ENSEMBLE_PRESETS = ("jet", "cavity", "obstacle")


def ensemble_specs(n=256, seed=4, sides=ENSEMBLE_SIDES, strokes=4):
    """Member descriptions of the ensemble: width and height uniform over `sides`, a
    preset uniform over jet / cavity / obstacle, and `strokes` solid Bresenham strokes
    (drawLine, main.go:809) from a uniform interior cell to a point at most a quarter of
    the shorter side away. Returns [dict(w, h, preset, strokes=[(x0,y0,x1,y1)])]."""
    rng = Rng(seed)
    out = []
    for _ in range(n):
        w = sides[rng.randint(0, len(sides) - 1)]
        h = sides[rng.randint(0, len(sides) - 1)]
        preset = ENSEMBLE_PRESETS[rng.randint(0, 2)]
        reach = max(1, min(w, h) // 4)
        st = []
        for _ in range(strokes):
            x0, y0 = rng.randint(1, w), rng.randint(1, h)
            x1 = min(max(x0 + rng.randint(-reach, reach), 1), w)
            y1 = min(max(y0 + rng.randint(-reach, reach), 1), h)
            st.append((x0, y0, x1, y1))
        out.append(dict(w=w, h=h, preset=preset, strokes=st))
    return out


def ensemble_member_setup(f, spec, iters=60):
    """One member: the preset (resetToPreset, main.go:747-793) with the jet of the jet
    scene (16 rows centred at i = 1, u = 1, smoke 1) or, for the obstacle, an inflow of
    u = 1 on the left edge as in the Karman scene; then the wall strokes. SL advection,
    no confinement."""
    if spec["preset"] == "jet":
        f.reset_to_preset(0)
        lo = f.NumY // 2 - 8
        f.set_jet(True, 1, max(lo, 1), min(lo + 16, f.NumY - 1), 1.0, 1.0)
    elif spec["preset"] == "cavity":
        f.reset_to_preset(1, 1.0)
    else:
        f.reset_to_preset(2)
        for j in range(1, f.NumY - 1):
            f.add_sink(1, j, 1.0, 0.0)
    for (x0, y0, x1, y1) in spec["strokes"]:
        f.draw_line(x0, y0, x1, y1, True)
    ensemble_knobs(f, iters)


def ensemble_knobs(f, iters=60):
    """The ensemble's solver knobs (on a member, a solo Fluid or the batch)."""
    f.UseBFECC = False
    f.Confinement = 0.0
    f.PressureIters = iters


def ensemble_impulses(step, specs, seed=4, sigma=8.0, fmax=5.0, every=10):
    """Gaussian impulses of a step: on every `every`-th step each member gets one at a
    uniform interior cell with |F| uniform in [0, fmax) and a uniform direction.
    Returns [(member, cx, cy, fx, fy, sigma)] (fx, fy rounded to float32)."""
    if step % every:
        return []
    rng = Rng(seed * 1000003 + step)
    out = []
    for e, s in enumerate(specs):
        cx, cy = rng.randint(1, s["w"]), rng.randint(1, s["h"])
        mag, ang = fmax * rng.uniform(), 2.0 * np.pi * rng.uniform()
        fx, fy = np.float32(mag * np.cos(ang)), np.float32(mag * np.sin(ang))
        out.append((e, cx, cy, float(fx), float(fy), sigma))
    return out
