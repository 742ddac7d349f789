"""Python mirror of the Go API of github.com/TheFellow/fluid/pkg/fluid over libfluidhip.

Method and field names follow the reference (New, Simulate, SetSolid, IsSolid,
SetVelocity, AddSmoke, Reset, ApplyForce, ApplyForceRadius, SetCircularObstacle,
GetAdaptiveTimeStep, MaxDivergence, SampleVelocity, Vorticity, VelocityMagnitude,
Pressure, Smoke, Velocity; knobs Confinement ... UseBFECC; package var Relaxation).
Out-of-range indices raise IndexError with the reference's panic text
(walls.go:6-11). The scene helpers restate package main's presets, jet, sources and
strokes (main/main.go), which the reference keeps outside pkg/fluid.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import FlParams, FlStats, FlTiming, check

FIELDS = {"U": 0, "V": 1, "newU": 2, "newV": 3, "p": 4, "S": 5, "M": 6, "newM": 7}
PASSES = {
    "copyBorderUV": 0,
    "applyViscosity": 1,
    "makeIncompressible": 2,
    "applyVorticityConfinement": 3,
    "addTurbulence": 4,
    "handleBorders": 5,
    "advectVelocity": 6,
    "advectSmoke": 7,
    "advectVelocityBFECC": 8,
    "advectSmokeBFECC": 9,
    "clearP": 10,
}
PRESETS = {"jet": 0, "cavity": 1, "karman": 2}

# Package variable `Relaxation` (fluid.go:7-9): read by every Simulate.
Relaxation = 1.9


class ScalarField:
    """ScalarField (scalar_field.go:7-22)."""

    def __init__(self, nx, ny, values, mn, mx):
        self.NumX, self.NumY = nx, ny
        self.values = values
        self.MinValue, self.MaxValue = mn, mx

    def Value(self, i, j):
        if i < 0 or i >= self.NumX:
            return 0.0, ValueError("x index (%d) out of range, must be between 0 and %d" % (i, self.NumX - 1))
        if j < 0 or j >= self.NumY:
            return 0.0, ValueError("y index (%d) out of range, must be between 0 and %d" % (j, self.NumY - 1))
        return float(self.values.reshape(-1)[i * self.NumY + j]), None


class VectorField:
    """VectorField (vector_field.go:5-20)."""

    def __init__(self, nx, ny, u, v):
        self.NumX, self.NumY = nx, ny
        self.valuesU, self.valuesV = u, v

    def Value(self, i, j):
        if i < 0 or i >= self.NumX:
            return 0.0, 0.0, ValueError("x index out of range, must be between 0 and %d" % (self.NumX - 1))
        if j < 0 or j >= self.NumY:
            return 0.0, 0.0, ValueError("y index out of range, must be between 0 and %d" % (self.NumY - 1))
        k = i * self.NumY + j
        return float(self.valuesU.reshape(-1)[k]), float(self.valuesV.reshape(-1)[k]), None


def _init_knobs(o):
    """The exported knobs of Fluid (fluid.go:27-39) with New's defaults (fluid.go:59-66)."""
    o.Confinement = 0.0
    o.ViscosityDiffusion = 0.0
    o.PressureDamping = 1.0
    o.TurbulenceStrength = 0.02
    o.SmokeAdvection = 1.0
    o.UseMultigrid = False
    o.MultigridLevels = 2
    o.UseBFECC = False
    # Additions: the reference hard-codes 8 sweeps and the early exit (fluid.go:81,175).
    o.PressureIters = 8
    o.EarlyExit = True
    o.Relaxation = None  # None: follow the module-level Relaxation like the Go package var
    o.Solver = 0         # 0 auto, 1 hyperplane kernel, 2 temporal-blocked kernel
    o.BlockSweeps = 0
    # Addition: back-trace in cell units (no reference counterpart; off = the reference's
    # absolute-position arithmetic, bit for bit).
    o.PreciseAdvection = False
    o.HandoffLag = 0     # temporal-blocked solve: 0 auto, 1 on time, 2 one block late
    o.BlockCols = 0      # temporal-blocked solve: columns per block, 0 auto (8), 4, 8 or 16


def _params_of(o):
    p = FlParams()
    p.confinement = o.Confinement
    p.viscosity_diffusion = o.ViscosityDiffusion
    p.pressure_damping = o.PressureDamping
    p.turbulence_strength = o.TurbulenceStrength
    p.smoke_advection = o.SmokeAdvection
    p.relaxation = Relaxation if o.Relaxation is None else o.Relaxation
    p.use_multigrid = int(bool(o.UseMultigrid))
    p.multigrid_levels = int(o.MultigridLevels)
    p.use_bfecc = int(bool(o.UseBFECC))
    p.pressure_iters = int(o.PressureIters)
    p.early_exit = int(bool(o.EarlyExit))
    p.solver = int(o.Solver)
    p.block_sweeps = int(o.BlockSweeps)
    p.precise_advection = int(bool(o.PreciseAdvection))
    p.handoff_lag = int(o.HandoffLag)
    p.block_cols = int(o.BlockCols)
    return p


class Fluid:
    """Fluid (fluid.go:11-40) backed by device buffers on one MI355X."""

    def __init__(self, density, width, height, h, _member=None):
        self._L = _lib.load()
        if _member is None:
            ctx = ctypes.c_void_p()
            check(self._L.fl_create(density, width, height, h, ctypes.byref(ctx)))
            self._owned = True
        else:
            ctx = _member
            self._owned = False  # a batch member belongs to its batch
        self._ctx = ctx
        nx, ny = ctypes.c_int(), ctypes.c_int()
        check(self._L.fl_dims(ctx, ctypes.byref(nx), ctypes.byref(ny)))
        self.NumX, self.NumY = nx.value, ny.value
        self._hval = float(np.float32(h))
        _init_knobs(self)

    @classmethod
    def New(cls, density, width, height, h):
        """New (fluid.go:42-68)."""
        return cls(density, width, height, h)

    def close(self):
        if getattr(self, "_ctx", None):
            if self._owned:
                self._L.fl_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def H(self):
        return self._hval

    def _push(self):
        p = _params_of(self)
        check(self._L.fl_set_params(self._ctx, ctypes.byref(p)))

    def autotune(self, dt):
        """Time the candidate Gauss-Seidel tile shapes (sweeps per tile, columns per block,
        handoff lag) on the current state; keep the fastest. Returns (sweeps per tile, ms)."""
        self._push()
        tk, ms = ctypes.c_int(), ctypes.c_float()
        check(self._L.fl_autotune(self._ctx, dt, ctypes.byref(tk), ctypes.byref(ms)))
        p = FlParams()
        check(self._L.fl_get_params(self._ctx, ctypes.byref(p)))
        self.BlockSweeps, self.BlockCols, self.HandoffLag = p.block_sweeps, p.block_cols, p.handoff_lag
        return tk.value, ms.value

    # ---- whole-field access ---------------------------------------------------------
    def get(self, name):
        a = np.empty((self.NumX, self.NumY), dtype=np.float32)
        check(self._L.fl_download(self._ctx, FIELDS[name], a.ctypes.data, a.size))
        return a

    def set(self, name, arr):
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.float32).reshape(self.NumX, self.NumY))
        check(self._L.fl_upload(self._ctx, FIELDS[name], a.ctypes.data, a.size))

    def fill_cells(self, name, i0, i1, j0, j1, value):
        """Direct writes name[i, j] = value for i in [i0, i1), j in [j0, j1) (any field but S),
        on the device."""
        check(self._L.fl_fill_cells(self._ctx, FIELDS[name], i0, i1, j0, j1, value))

    def state(self):
        return {k: self.get(k) for k in FIELDS}

    def load_state(self, st):
        for k, v in st.items():
            self.set(k, v)

    # The exported slices U, V, S, M (fluid.go:17-22) as downloaded copies.
    U = property(lambda self: self.get("U"))
    V = property(lambda self: self.get("V"))
    S = property(lambda self: self.get("S"))
    M = property(lambda self: self.get("M"))

    # ---- step -------------------------------------------------------------------------
    def Simulate(self, dt):
        """Simulate (fluid.go:79-109)."""
        self._push()
        check(self._L.fl_simulate(self._ctx, dt))

    def sync(self):
        check(self._L.fl_sync(self._ctx))

    def stats(self):
        s = FlStats()
        check(self._L.fl_get_stats(self._ctx, ctypes.byref(s)))
        return {"sweeps_run": s.sweeps_run, "reran": bool(s.reran), "invalid_sample": bool(s.invalid_sample),
                "max_div_last": s.max_div_last}

    def set_timing(self, on):
        """Per-pass device timing with HIP events on the context's stream."""
        check(self._L.fl_set_timing(self._ctx, int(bool(on))))

    def timing(self):
        t = FlTiming()
        check(self._L.fl_get_timing(self._ctx, ctypes.byref(t)))
        return {"total_ms": t.total_ms, "pressure_ms": t.pressure_ms, "advect_ms": t.advect_ms,
                "other_ms": t.other_ms, "steps": t.steps}

    def sweep_maxdiv(self):
        cnt = ctypes.c_int()
        check(self._L.fl_sweep_maxdiv(self._ctx, None, 0, ctypes.byref(cnt)))
        out = np.zeros(cnt.value, dtype=np.float32)
        if cnt.value:
            check(self._L.fl_sweep_maxdiv(self._ctx, out.ctypes.data, cnt.value, ctypes.byref(cnt)))
        return out

    def gs_layout(self):
        """The tiling of the last temporal-blocked solve (fl_get_gs_layout) as a dict."""
        out = _lib.FlGsLayout()
        check(self._L.fl_get_gs_layout(self._ctx, ctypes.byref(out)))
        return {k: getattr(out, k) for k, _ in _lib.FlGsLayout._fields_}

    def run_pass(self, name, dt=0.0, iters=0):
        """Run one unexported pass (copyBorder, makeIncompressible, advectVelocity, ...)."""
        self._push()
        check(self._L.fl_pass(self._ctx, PASSES[name], dt, iters))

    # ---- walls.go -----------------------------------------------------------------------
    def last_error(self):
        """The text of the last failing call on this context (fl_ctx_last_error)."""
        buf = ctypes.create_string_buffer(512)
        self._L.fl_ctx_last_error(self._ctx, buf, len(buf))
        return buf.value.decode()

    def _idx(self, rc):
        if rc in (2, 3):
            raise IndexError(self.last_error())
        check(rc)

    def SetSolid(self, i, j, value):
        self._idx(self._L.fl_set_solid(self._ctx, i, j, int(bool(value))))

    def IsSolid(self, i, j):
        out = ctypes.c_int()
        self._idx(self._L.fl_is_solid(self._ctx, i, j, ctypes.byref(out)))
        return bool(out.value)

    def SetVelocity(self, i, j, u, v):
        self._idx(self._L.fl_set_velocity(self._ctx, i, j, u, v))

    def AddSmoke(self, i, j, s):
        self._idx(self._L.fl_add_smoke(self._ctx, i, j, s))

    def Reset(self):
        check(self._L.fl_reset(self._ctx))

    # ---- fluid.go API ---------------------------------------------------------------------
    def ApplyForce(self, i, j, fx, fy):
        check(self._L.fl_apply_force(self._ctx, i, j, fx, fy))

    def ApplyForceRadius(self, cx, cy, fx, fy, radius):
        check(self._L.fl_apply_force_radius(self._ctx, cx, cy, fx, fy, radius))

    def ApplyForceGaussian(self, cx, cy, fx, fy, sigma):
        """ApplyForce (fluid.go:761) weighted by a Gaussian of width sigma, truncated at 3 sigma."""
        check(self._L.fl_apply_force_gaussian(self._ctx, cx, cy, fx, fy, sigma))

    def render(self, view, overlay=True):
        """Game.Draw's RGBA image (NumY x NumX x 4): view 0 smoke, 1 pressure, 2 speed, 3 vorticity."""
        pix = np.zeros((self.NumY, self.NumX, 4), dtype=np.uint8)
        check(self._L.fl_render(self._ctx, int(view), int(bool(overlay)), pix.ctypes.data, pix.size))
        return pix

    def advect_particles(self, dt, x, y, age, max_age, no_alloc=False):
        """Game.advectParticles (main.go:512-546) on float32 arrays, in place; returns keep.
        no_alloc=True stages the particles in the library's static device array
        (fl_advect_particles_noalloc: at most PARTICLES_NOALLOC_MAX particles, no allocation)."""
        arrs = (x, y, age, max_age)
        for a in arrs:
            if not isinstance(a, np.ndarray) or a.dtype != np.float32 or not a.flags.c_contiguous:
                raise TypeError("particle arrays must be C-contiguous float32 numpy arrays")
            if a.size != x.size:
                raise ValueError("particle arrays must have the same length")
        keep = np.zeros(x.size, dtype=np.uint8)
        fn = self._L.fl_advect_particles_noalloc if no_alloc else self._L.fl_advect_particles
        check(fn(self._ctx, dt, x.size, x.ctypes.data, y.ctypes.data, age.ctypes.data, max_age.ctypes.data,
                 keep.ctypes.data))
        return keep

    def SetCircularObstacle(self, cx, cy, radius):
        check(self._L.fl_set_circular_obstacle(self._ctx, cx, cy, radius))

    def GetAdaptiveTimeStep(self, basedt):
        out = ctypes.c_float()
        self._push()
        check(self._L.fl_get_adaptive_time_step(self._ctx, basedt, ctypes.byref(out)))
        return out.value

    def MaxDivergence(self):
        out = ctypes.c_float()
        check(self._L.fl_max_divergence(self._ctx, ctypes.byref(out)))
        return out.value

    def SampleVelocity(self, x, y):
        u, v = ctypes.c_float(), ctypes.c_float()
        check(self._L.fl_sample_velocity(self._ctx, x, y, ctypes.byref(u), ctypes.byref(v)))
        return u.value, v.value

    def SampleVelocityBatch(self, xs, ys):
        xs = np.ascontiguousarray(xs, dtype=np.float32).reshape(-1)
        ys = np.ascontiguousarray(ys, dtype=np.float32).reshape(-1)
        if xs.size != ys.size:
            raise ValueError("xs and ys must have the same length (%d != %d)" % (xs.size, ys.size))
        us, vs = np.empty_like(xs), np.empty_like(ys)
        check(self._L.fl_sample_velocity_batch(self._ctx, xs.ctypes.data, ys.ctypes.data, xs.size, us.ctypes.data,
                                               vs.ctypes.data))
        return us, vs

    def _view(self, fn):
        out = np.empty((self.NumX, self.NumY), dtype=np.float32)
        mn, mx = ctypes.c_float(), ctypes.c_float()
        check(fn(self._ctx, out.ctypes.data, out.size, ctypes.byref(mn), ctypes.byref(mx)))
        return ScalarField(self.NumX, self.NumY, out, mn.value, mx.value)

    def Vorticity(self):
        return self._view(self._L.fl_vorticity)

    def VelocityMagnitude(self):
        return self._view(self._L.fl_velocity_magnitude)

    def _minmax(self, name):
        mn, mx = ctypes.c_float(), ctypes.c_float()
        check(self._L.fl_field_minmax(self._ctx, FIELDS[name], ctypes.byref(mn), ctypes.byref(mx)))
        return ScalarField(self.NumX, self.NumY, self.get(name), mn.value, mx.value)

    def field_range(self, name):
        """(MinValue, MaxValue) of a field, reduced on the device; nothing is downloaded."""
        mn, mx = ctypes.c_float(), ctypes.c_float()
        check(self._L.fl_field_minmax(self._ctx, FIELDS[name], ctypes.byref(mn), ctypes.byref(mx)))
        return mn.value, mx.value

    def Pressure(self):
        return self._minmax("p")

    def Smoke(self):
        return self._minmax("M")

    def Velocity(self):
        return VectorField(self.NumX, self.NumY, self.get("U"), self.get("V"))

    # ---- scene layer (package main) ----------------------------------------------------------
    def scene_walls(self, top, bottom, left, right):
        check(self._L.fl_scene_set_walls(self._ctx, int(top), int(bottom), int(left), int(right)))

    def apply_wall_settings(self):
        check(self._L.fl_scene_apply_walls(self._ctx))

    def clear_walls(self):
        check(self._L.fl_scene_clear_walls(self._ctx))

    def reset_to_preset(self, preset, lid_u=3.0):
        check(self._L.fl_scene_preset(self._ctx, PRESETS.get(preset, preset), lid_u))

    def set_jet(self, on, x=1, lo=0, hi=0, u=4.0, smoke=1.0):
        self._idx(self._L.fl_scene_set_jet(self._ctx, int(bool(on)), x, lo, hi, u, smoke))

    def add_source(self, i, j, u, v):
        check(self._L.fl_scene_add_source(self._ctx, 0, i, j, u, v))

    def add_sink(self, i, j, u, v):
        check(self._L.fl_scene_add_source(self._ctx, 1, i, j, u, v))

    def clear_sources(self):
        check(self._L.fl_scene_clear_sources(self._ctx))

    def apply_sources(self):
        check(self._L.fl_scene_apply_sources(self._ctx))

    def update(self, dt):
        """Game.Update minus input and particles (main.go:231-249): sources, jet, Simulate."""
        self._push()
        check(self._L.fl_scene_update(self._ctx, dt))

    def draw_line(self, x0, y0, x1, y1, value):
        check(self._L.fl_scene_draw_line(self._ctx, x0, y0, x1, y1, int(bool(value))))

    def smoke_brush(self, ci, cj, r, strength):
        check(self._L.fl_scene_smoke_brush(self._ctx, ci, cj, r, strength))


def New(density, width, height, h):
    """New (fluid.go:42)."""
    return Fluid(density, width, height, h)


class FluidBatch:
    """An ensemble of independent Fluids of uneven sizes advanced together.

    In the reference an ensemble is a loop of Simulate calls, one per Fluid (fluid.go:79,
    main.go:231-249 per frame). Here `update(dt)` steps every member with one launch per
    pass and one persistent pressure launch over all members' tiles, bit-identically to
    calling `update(dt)` (sources, jet, Simulate) on each member alone with the batch's
    knobs. `members[e]` is a Fluid view for setup, edits, scenes and downloads.
    """

    def __init__(self, density, sizes, h):
        self._L = _lib.load()
        n = len(sizes)
        ws = (ctypes.c_int * n)(*[int(w) for w, _ in sizes])
        hs = (ctypes.c_int * n)(*[int(hh) for _, hh in sizes])
        b = ctypes.c_void_p()
        check(self._L.fl_batch_create(n, ws, hs, density, h, ctypes.byref(b)))
        self._b = b
        self.members = []
        for e in range(n):
            m = ctypes.c_void_p()
            check(self._L.fl_batch_member(b, e, ctypes.byref(m)))
            self.members.append(Fluid(density, sizes[e][0], sizes[e][1], h, _member=m))
        _init_knobs(self)

    def close(self):
        if getattr(self, "_b", None):
            for m in self.members:
                m._ctx = None
            self._L.fl_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.members)

    def _push(self):
        p = _params_of(self)
        check(self._L.fl_batch_set_params(self._b, ctypes.byref(p)))

    def update(self, dt):
        self._push()
        check(self._L.fl_batch_update(self._b, dt))

    def sync(self):
        check(self._L.fl_batch_sync(self._b))

    def set_graph(self, on):
        check(self._L.fl_batch_set_graph(self._b, int(bool(on))))

    def autotune(self, dt):
        """fl_batch_autotune: the batched solve's tile shape, timed on the live state."""
        self._push()
        tk, ms = ctypes.c_int(), ctypes.c_float()
        check(self._L.fl_batch_autotune(self._b, dt, ctypes.byref(tk), ctypes.byref(ms)))
        p = FlParams()
        check(self._L.fl_batch_get_params(self._b, ctypes.byref(p)))
        self.BlockSweeps, self.BlockCols, self.HandoffLag = p.block_sweeps, p.block_cols, p.handoff_lag
        return tk.value, ms.value


def device_count():
    n = ctypes.c_int()
    _lib.load().fl_device_count(ctypes.byref(n))
    return n.value


class Rollout:
    """A differentiable rollout of `steps` steps from the Fluid's current state.

    Each step is the scene update (sources and jet, main.go:231-241), then U,V += F on fluid
    interior cells, then Simulate (fluid.go:79-109); the loss is sum((M - target)^2) after
    the last step. `forward` restores the captured initial state, runs the steps on the
    device (checkpointing every `checkpoint_every` steps) and returns the loss; `backward`
    returns (dL/dFu, dL/dFv) from the hand-written adjoint. The reference has no gradients.
    """

    def __init__(self, fluid, steps, checkpoint_every=8):
        self._L = fluid._L
        self._f = fluid
        v = ctypes.c_void_p()
        fluid._push()
        check(self._L.fl_vjp_create(fluid._ctx, int(steps), int(checkpoint_every), ctypes.byref(v)))
        self._v = v
        self.steps = steps
        self.shape = (fluid.NumX, fluid.NumY)

    def close(self):
        if getattr(self, "_v", None):
            self._L.fl_vjp_destroy(self._v)
            self._v = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _arr(self, a):
        return np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(self.shape))

    def forward(self, dt, Fu, Fv, target):
        self._f._push()
        fu, fv, tg = self._arr(Fu), self._arr(Fv), self._arr(target)
        loss = ctypes.c_double()
        fp = ctypes.POINTER(ctypes.c_float)
        check(self._L.fl_vjp_forward(self._v, dt, fu.ctypes.data_as(fp), fv.ctypes.data_as(fp),
                                     tg.ctypes.data_as(fp), fu.size, ctypes.byref(loss)))
        return loss.value

    def backward(self):
        gu = np.zeros(self.shape, dtype=np.float32)
        gv = np.zeros(self.shape, dtype=np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        check(self._L.fl_vjp_backward(self._v, gu.ctypes.data_as(fp), gv.ctypes.data_as(fp), gu.size))
        return gu, gv
