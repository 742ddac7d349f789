"""fl_advect_particles_noalloc (static device scratch, no allocation) against
fl_advect_particles, through the C ABI: bit for bit on two calls in a row, the second with
fewer particles, plus the size limit."""
import ctypes

import numpy as np
import pytest

from fluid_amd import Fluid, FluidError
from fluid_amd import _lib
from fluid_amd import configs as C

pytestmark = pytest.mark.gpu


def particles(rng, n):
    h = np.float32(C.H)
    x = (rng.random(n) * 100 - 1).astype(np.float32) * h   # some start outside the grid
    y = (rng.random(n) * 76 - 1).astype(np.float32) * h
    age = (rng.random(n) * 1.4).astype(np.float32)
    mx = np.full(n, 1.5, np.float32)
    return x, y, age, mx


def run_both(f, arrs):
    a = [v.copy() for v in arrs]
    b = [v.copy() for v in arrs]
    ka = f.advect_particles(np.float32(C.DT), *a)
    kb = f.advect_particles(np.float32(C.DT), *b, no_alloc=True)
    assert np.array_equal(ka, kb)
    for u, v in zip(a[:3], b[:3]):
        assert np.array_equal(u.view(np.uint32), v.view(np.uint32))
    return ka


def test_noalloc_particles_equal_the_allocating_call():
    print("no-alloc particles: jet 96x72, 9000 then 700 particles", flush=True)
    f = C.jet_scene(lambda *a: Fluid(*a), 96, 72, 20, jet_rows=12)
    f.SetCircularObstacle(30, 36, 8)
    for _ in range(4):
        f.update(C.DT)
    rng = np.random.default_rng(21)
    keep = run_both(f, particles(rng, 9000))
    assert 0 < keep.sum() < keep.size
    f.update(C.DT)
    keep = run_both(f, particles(rng, 700))  # smaller: the tail of the first call stays in the array
    assert 0 < keep.sum() < keep.size
    # at the limit, and queued edits flushed first
    f.SetVelocity(40, 30, 2.0, -1.0)
    run_both(f, particles(rng, _lib.PARTICLES_NOALLOC_MAX))


def test_noalloc_particles_reject_too_many():
    f = C.jet_scene(lambda *a: Fluid(*a), 32, 24, 8, jet_rows=6)
    n = _lib.PARTICLES_NOALLOC_MAX + 1
    x, y, age, mx = particles(np.random.default_rng(3), n)
    keep = np.zeros(n, np.uint8)
    rc = f._L.fl_advect_particles_noalloc(f._ctx, ctypes.c_float(C.DT), n, x.ctypes.data, y.ctypes.data,
                                          age.ctypes.data, mx.ctypes.data, keep.ctypes.data)
    assert rc == 1
    assert b"exceeds the static scratch" in f._L.fl_last_error()
    with pytest.raises(FluidError):
        f.advect_particles(np.float32(C.DT), x, y, age, mx, no_alloc=True)
