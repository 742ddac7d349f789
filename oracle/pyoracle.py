"""ctypes wrapper of the CPU oracle (oracle/fluid_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker. The product (fluid_amd) never imports it.

`OracleFluid` exposes the same method names as fluid_amd.Fluid (which mirror the Go
API of pkg/fluid), so scene builders and tests drive either object the same way.
"""
import contextlib
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle_fluid.so")
PLANT_LIB_PATH = os.path.join(HERE, "_build", "liboracle_fluid_plant.so")

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

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load(path):
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    _bind(L)
    return L


def lib():
    global _lib
    if _lib is None:
        _lib = _load(LIB_PATH)
    return _lib


@contextlib.contextmanager
def planted(k):
    """Run the block on the plantable build of the oracle with planted mistake k switched on
    (the list is at the top of fluid_oracle.cpp). Objects made inside use that build."""
    global _lib
    L = _load(PLANT_LIB_PATH)
    L.orc_set_plant.argtypes = [ctypes.c_int]
    L.orc_set_plant(int(k))
    saved, _lib = _lib, L
    try:
        yield
    finally:
        L.orc_set_plant(0)
        _lib = saved


def _bind(L):
    vp, i, f, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_double
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    sig = {
        "orc_go_sin": (d, [d]),
        "orc_turb_noise": (f, [i]),
        "orc_noise_table": (None, [i, vp]),
        "orc_go_sin_many": (None, [i, vp, vp]),
        "orc_new": (vp, [i, f, i, i, f]),
        "orc_free": (None, [vp]),
        "orc_nx": (i, [vp]),
        "orc_ny": (i, [vp]),
        "orc_set_field": (i, [vp, i, vp]),
        "orc_get_field": (i, [vp, i, vp]),
        "orc_set_params": (None, [vp, dp, ip]),
        "orc_set_threads": (None, [vp, i]),
        "orc_simulate": (None, [vp, f]),
        "orc_pass": (i, [vp, i, f, i]),
        "orc_last_sweeps": (i, [vp]),
        "orc_sweep_maxdiv": (i, [vp, dp, i]),
        "orc_invalid_index": (i, [vp]),
        "orc_set_forced_sweeps": (None, [vp, i]),
        "orc_get_turb_gate": (None, [vp, vp]),
        "orc_set_turb_gate": (None, [vp, vp]),
        "orc_set_solid": (i, [vp, i, i, i]),
        "orc_is_solid": (i, [vp, i, i, ip]),
        "orc_set_velocity": (i, [vp, i, i, d, d]),
        "orc_add_smoke": (i, [vp, i, i, d]),
        "orc_reset": (None, [vp]),
        "orc_apply_force": (None, [vp, i, i, d, d]),
        "orc_apply_force_radius": (None, [vp, i, i, d, d, i]),
        "orc_apply_force_gaussian": (None, [vp, i, i, d, d, d]),
        "orc_set_circular_obstacle": (None, [vp, i, i, i]),
        "orc_adaptive_dt": (d, [vp, f]),
        "orc_max_divergence": (d, [vp]),
        "orc_sample_velocity": (None, [vp, f, f, dp, dp]),
        "orc_view": (None, [vp, i, vp, dp, dp]),
        "orc_field_minmax": (None, [vp, i, dp, dp]),
        "orc_scene_walls": (None, [vp, i, i, i, i]),
        "orc_scene_apply_walls": (None, [vp]),
        "orc_scene_clear_walls": (None, [vp]),
        "orc_scene_preset": (None, [vp, i, f]),
        "orc_scene_jet": (None, [vp, i, i, i, i, f, f]),
        "orc_scene_jet_on": (i, [vp]),
        "orc_scene_add_source": (None, [vp, i, i, i, f, f]),
        "orc_scene_clear_sources": (None, [vp]),
        "orc_scene_num_sources": (i, [vp, i]),
        "orc_scene_apply_sources": (None, [vp]),
        "orc_scene_update": (None, [vp, f]),
        "orc_scene_draw_line": (None, [vp, i, i, i, i, i]),
        "orc_scene_smoke_brush": (None, [vp, i, i, i, f]),
        "orc_rollout": (d, [vp, i, f, vp, vp, vp, vp, vp]),
        "orc_advect_particles": (None, [vp, f, i, vp, vp, vp, vp, vp]),
        "orc_render": (None, [vp, i, i, vp]),
        "orc_sample_field": (d, [vp, i, i, d, d]),
        "orc_mg_dims": (None, [vp, ip, ip]),
        "orc_mg_residual": (None, [vp, vp]),
        "orc_mg_restrict": (None, [vp, vp, vp]),
        "orc_mg_coarse_solve": (None, [vp, vp, vp]),
        "orc_mg_prolong": (None, [vp, vp, vp]),
        "orc_mg_correct": (None, [vp, vp, d]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def go_sin(x):
    return lib().orc_go_sin(float(x))


def turb_noise(k):
    return lib().orc_turb_noise(int(k))


def noise_table(n):
    out = np.empty(n, dtype=np.float32)
    lib().orc_noise_table(int(n), out.ctypes.data)
    return out


def go_sin_many(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    lib().orc_go_sin_many(x.size, x.ctypes.data, out.ctypes.data)
    return out


class ScalarField:
    """pkg/fluid ScalarField (scalar_field.go): values, NumX, NumY, MinValue, MaxValue."""

    def __init__(self, nx, ny, values, mn, mx):
        self.NumX, self.NumY = nx, ny
        self.values = values
        self.MinValue, self.MaxValue = mn, mx

    def Value(self, i, j):
        if i < 0 or i >= self.NumX:
            return 0.0, ValueError("x index (%d) out of range, must be between 0 and %d" % (i, self.NumX - 1))
        if j < 0 or j >= self.NumY:
            return 0.0, ValueError("y index (%d) out of range, must be between 0 and %d" % (j, self.NumY - 1))
        return self.values.reshape(-1)[i * self.NumY + j], None


class OracleFluid:
    """CPU restatement of pkg/fluid's Fluid; prec=32 is float32-faithful, 64 is the FP64 model."""

    KNOBS = ("Confinement", "ViscosityDiffusion", "PressureDamping", "TurbulenceStrength", "SmokeAdvection",
             "Relaxation", "UseMultigrid", "MultigridLevels", "UseBFECC", "PressureIters", "EarlyExit",
             "PreciseAdvection")

    def __init__(self, density, width, height, h, prec=32, threads=1):
        self._L = lib()
        self.prec = prec
        self.dtype = np.float32 if prec == 32 else np.float64
        self._h = self._L.orc_new(prec, density, width, height, h)
        if not self._h:
            raise ValueError("bad oracle arguments")
        self.NumX = self._L.orc_nx(self._h)
        self.NumY = self._L.orc_ny(self._h)
        self.hval = float(np.float32(h))
        self.Confinement = 0.0
        self.ViscosityDiffusion = 0.0
        self.PressureDamping = 1.0
        self.TurbulenceStrength = 0.02
        self.SmokeAdvection = 1.0
        self.Relaxation = 1.9
        self.UseMultigrid = False
        self.MultigridLevels = 2
        self.UseBFECC = False
        self.PressureIters = 8
        self.EarlyExit = True
        self.PreciseAdvection = False
        self._L.orc_set_threads(self._h, threads)

    New = classmethod(lambda cls, density, width, height, h, **kw: cls(density, width, height, h, **kw))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.orc_free(h)
            self._h = None

    def H(self):
        return self.hval

    def set_threads(self, n):
        self._L.orc_set_threads(self._h, int(n))

    def _push(self):
        fp = (ctypes.c_double * 6)(self.Confinement, self.ViscosityDiffusion, self.PressureDamping,
                                   self.TurbulenceStrength, self.SmokeAdvection, self.Relaxation)
        ip = (ctypes.c_int * 6)(int(self.UseMultigrid), int(self.MultigridLevels), int(self.UseBFECC),
                                int(self.PressureIters), int(self.EarlyExit), int(self.PreciseAdvection))
        self._L.orc_set_params(self._h, fp, ip)

    # ---- particles and rendering (float32 handles) ----
    def advect_particles(self, dt, x, y, age, max_age):
        """Game.advectParticles (main.go:512-546) on float32 arrays, in place; returns keep."""
        assert self.prec == 32
        for a in (x, y, age, max_age):
            assert a.dtype == np.float32 and a.flags.c_contiguous
        keep = np.zeros(x.size, dtype=np.uint8)
        self._L.orc_advect_particles(self._h, dt, x.size, x.ctypes.data, y.ctypes.data, age.ctypes.data,
                                     max_age.ctypes.data, keep.ctypes.data)
        return keep

    def render(self, view, overlay=True):
        """Game.Draw's RGBA image (NumY rows x NumX columns x 4), main.go:549-652."""
        assert self.prec == 32
        pix = np.zeros((self.NumY, self.NumX, 4), dtype=np.uint8)
        self._L.orc_render(self._h, int(view), int(bool(overlay)), pix.ctypes.data)
        return pix

    # ---- differentiable rollout (oracle/fluid_adjoint.inc) ----
    def rollout(self, steps, dt, Fu, Fv, target, grad=True):
        """Loss sum((M - target)^2) after `steps` steps of sources+jet, U,V += F on fluid
        interior cells, Simulate; with grad, also (dL/dFu, dL/dFv). The state is kept."""
        self._push()
        shape = (self.NumX, self.NumY)
        fu = np.ascontiguousarray(np.asarray(Fu, dtype=self.dtype).reshape(shape))
        fv = np.ascontiguousarray(np.asarray(Fv, dtype=self.dtype).reshape(shape))
        tg = np.ascontiguousarray(np.asarray(target, dtype=self.dtype).reshape(shape))
        if not grad:
            return self._L.orc_rollout(self._h, int(steps), dt, fu.ctypes.data, fv.ctypes.data, tg.ctypes.data,
                                       None, None)
        gu = np.zeros(shape, dtype=self.dtype)
        gv = np.zeros(shape, dtype=self.dtype)
        loss = self._L.orc_rollout(self._h, int(steps), dt, fu.ctypes.data, fv.ctypes.data, tg.ctypes.data,
                                   gu.ctypes.data, gv.ctypes.data)
        return loss, gu, gv

    # ---- fields ----
    def get(self, name):
        a = np.empty((self.NumX, self.NumY), dtype=self.dtype)
        self._L.orc_get_field(self._h, FIELDS[name], a.ctypes.data)
        return a

    def set(self, name, arr):
        a = np.ascontiguousarray(np.asarray(arr, dtype=self.dtype).reshape(self.NumX, self.NumY))
        self._L.orc_set_field(self._h, FIELDS[name], a.ctypes.data)

    def fill_cells(self, name, i0, i1, j0, j1, value):
        assert name != "S"
        a = self.get(name)
        a[i0:i1, j0:j1] = value
        self.set(name, a)

    def state(self):
        return {k: self.get(k) for k in FIELDS}

    def load_state(self, st):
        for k, v in st.items():
            self.set(k, v)

    # ---- step and passes ----
    def Simulate(self, dt):
        self._push()
        self._L.orc_simulate(self._h, dt)

    def run_pass(self, name, dt=0.0, iters=0):
        self._push()
        if self._L.orc_pass(self._h, PASSES[name], dt, iters):
            raise ValueError(name)

    def last_sweeps(self):
        return self._L.orc_last_sweeps(self._h)

    def sweep_maxdiv(self):
        buf = (ctypes.c_double * 100000)()
        n = self._L.orc_sweep_maxdiv(self._h, buf, 100000)
        return np.array(buf[:n])

    def invalid_index(self):
        return bool(self._L.orc_invalid_index(self._h))

    def set_forced_sweeps(self, n):
        self._L.orc_set_forced_sweeps(self._h, int(n))

    def turb_gate(self):
        g = np.empty(self.NumX * self.NumY, dtype=np.uint8)
        self._L.orc_get_turb_gate(self._h, g.ctypes.data)
        return g

    def set_turb_gate(self, gate):
        if gate is None:
            self._L.orc_set_turb_gate(self._h, None)
            self._gate = None
        else:
            self._gate = np.ascontiguousarray(gate, dtype=np.uint8)
            self._L.orc_set_turb_gate(self._h, self._gate.ctypes.data)

    # ---- walls.go ----
    def _chk(self, rc, i, j):
        if rc == 1:
            raise IndexError("invalid x-index: %d" % i)
        if rc == 2:
            raise IndexError("invalid y-index: %d" % j)

    def SetSolid(self, i, j, v):
        self._chk(self._L.orc_set_solid(self._h, i, j, int(bool(v))), i, j)

    def IsSolid(self, i, j):
        out = ctypes.c_int()
        self._chk(self._L.orc_is_solid(self._h, i, j, ctypes.byref(out)), i, j)
        return bool(out.value)

    def SetVelocity(self, i, j, u, v):
        self._chk(self._L.orc_set_velocity(self._h, i, j, u, v), i, j)

    def AddSmoke(self, i, j, s):
        self._chk(self._L.orc_add_smoke(self._h, i, j, s), i, j)

    def Reset(self):
        self._L.orc_reset(self._h)

    def ApplyForce(self, i, j, fx, fy):
        self._L.orc_apply_force(self._h, i, j, fx, fy)

    def ApplyForceRadius(self, cx, cy, fx, fy, r):
        self._L.orc_apply_force_radius(self._h, cx, cy, fx, fy, r)

    def ApplyForceGaussian(self, cx, cy, fx, fy, sigma):
        self._L.orc_apply_force_gaussian(self._h, cx, cy, fx, fy, sigma)

    def SetCircularObstacle(self, cx, cy, r):
        self._L.orc_set_circular_obstacle(self._h, cx, cy, r)

    # ---- single functions for the closed-form pins ----
    def sample_field(self, kind, name, x, y):
        """sampleField (fluid.go:357-398) of field `name` at (x, y); kind "U", "V" or "M"
        selects the half-cell offsets."""
        return self._L.orc_sample_field(self._h, {"U": 0, "V": 1, "M": 2}[kind], FIELDS[name], float(x),
                                        float(y))

    def mg_dims(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        self._L.orc_mg_dims(self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def mg_residual(self):
        """computePressureResidual (fluid.go:601-633) of the current state."""
        out = np.empty((self.NumX, self.NumY), dtype=self.dtype)
        self._L.orc_mg_residual(self._h, out.ctypes.data)
        return out

    def mg_restrict(self, fine):
        """restrictResidual (fluid.go:634-665)."""
        fine = np.ascontiguousarray(np.asarray(fine, dtype=self.dtype).reshape(self.NumX, self.NumY))
        out = np.empty(self.mg_dims(), dtype=self.dtype)
        self._L.orc_mg_restrict(self._h, fine.ctypes.data, out.ctypes.data)
        return out

    def mg_coarse_solve(self, rhs):
        """solveCoarseGrid (fluid.go:667-726): 40 GS sweeps at 1.6 on the coarse mask."""
        rhs = np.ascontiguousarray(np.asarray(rhs, dtype=self.dtype).reshape(self.mg_dims()))
        out = np.empty(self.mg_dims(), dtype=self.dtype)
        self._L.orc_mg_coarse_solve(self._h, rhs.ctypes.data, out.ctypes.data)
        return out

    def mg_prolong(self, coarse):
        """prolongateCorrection (fluid.go:728-758)."""
        coarse = np.ascontiguousarray(np.asarray(coarse, dtype=self.dtype).reshape(self.mg_dims()))
        out = np.empty((self.NumX, self.NumY), dtype=self.dtype)
        self._L.orc_mg_prolong(self._h, coarse.ctypes.data, out.ctypes.data)
        return out

    def mg_correct(self, corr, cp):
        """applePressureCorrection (fluid.go:1123-1149), in place on p, U, V."""
        corr = np.ascontiguousarray(np.asarray(corr, dtype=self.dtype).reshape(self.NumX, self.NumY))
        self._L.orc_mg_correct(self._h, corr.ctypes.data, float(cp))

    # ---- views ----
    def GetAdaptiveTimeStep(self, basedt):
        return self._L.orc_adaptive_dt(self._h, basedt)

    def MaxDivergence(self):
        return self._L.orc_max_divergence(self._h)

    def SampleVelocity(self, x, y):
        u, v = ctypes.c_double(), ctypes.c_double()
        self._L.orc_sample_velocity(self._h, x, y, ctypes.byref(u), ctypes.byref(v))
        return u.value, v.value

    def _view(self, kind):
        out = np.empty((self.NumX, self.NumY), dtype=self.dtype)
        mn, mx = ctypes.c_double(), ctypes.c_double()
        self._L.orc_view(self._h, kind, out.ctypes.data, ctypes.byref(mn), ctypes.byref(mx))
        return ScalarField(self.NumX, self.NumY, out, mn.value, mx.value)

    def Vorticity(self):
        return self._view(0)

    def VelocityMagnitude(self):
        return self._view(1)

    def _minmax_field(self, name):
        mn, mx = ctypes.c_double(), ctypes.c_double()
        self._L.orc_field_minmax(self._h, FIELDS[name], ctypes.byref(mn), ctypes.byref(mx))
        return ScalarField(self.NumX, self.NumY, self.get(name), mn.value, mx.value)

    def Pressure(self):
        return self._minmax_field("p")

    def Smoke(self):
        return self._minmax_field("M")

    # ---- scene (package main) ----
    def scene_walls(self, top, bottom, left, right):
        self._L.orc_scene_walls(self._h, int(top), int(bottom), int(left), int(right))

    def apply_wall_settings(self):
        self._L.orc_scene_apply_walls(self._h)

    def clear_walls(self):
        self._L.orc_scene_clear_walls(self._h)

    def reset_to_preset(self, preset, lid_u=3.0):
        self._L.orc_scene_preset(self._h, {"jet": 0, "cavity": 1, "karman": 2}.get(preset, preset), lid_u)

    def set_jet(self, on, x=1, lo=0, hi=0, u=4.0, smoke=1.0):
        self._L.orc_scene_jet(self._h, int(on), x, lo, hi, u, smoke)

    def add_source(self, i, j, u, v):
        self._L.orc_scene_add_source(self._h, 0, i, j, u, v)

    def add_sink(self, i, j, u, v):
        self._L.orc_scene_add_source(self._h, 1, i, j, u, v)

    def clear_sources(self):
        self._L.orc_scene_clear_sources(self._h)

    def num_sources(self, sink=False):
        return self._L.orc_scene_num_sources(self._h, int(sink))

    def apply_sources(self):
        self._L.orc_scene_apply_sources(self._h)

    def update(self, dt):
        self._push()
        self._L.orc_scene_update(self._h, dt)

    def draw_line(self, x0, y0, x1, y1, v):
        self._L.orc_scene_draw_line(self._h, x0, y0, x1, y1, int(bool(v)))

    def smoke_brush(self, ci, cj, r, strength):
        self._L.orc_scene_smoke_brush(self._h, ci, cj, r, strength)
