"""CPU checks of the drop-in boundary: the C-ABI library loads and exports every entry
point include/fluid_hip.h declares; without a GPU it fails loudly (no CPU fallback)."""
import ctypes
import os
import re

import numpy as np
import pytest

from fluid_amd import _lib
from fluid_amd.configs import Rng, cavity_discs, splitmix64

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "fluid_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fl_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared_symbols()
    assert len(names) >= 40
    for n in names:
        assert hasattr(L, n), n
    # the ctypes table covers the whole header, so the Python mirror binds every entry point
    assert set(names) == set(_lib.SIGNATURES)


def test_no_device_fails_loudly():
    L = _lib.load()
    n = ctypes.c_int(-1)
    assert L.fl_device_count(ctypes.byref(n)) == 0
    if n.value > 0:
        pytest.skip("a GPU is present")
    ctx = ctypes.c_void_p()
    rc = L.fl_create(1000.0, 16, 16, 0.01, ctypes.byref(ctx))
    assert rc == 4 and not ctx.value
    assert b"no HIP device" in L.fl_last_error()


def test_invalid_arguments_rejected_before_device():
    L = _lib.load()
    ctx = ctypes.c_void_p()
    assert L.fl_create(1000.0, 0, 16, 0.01, ctypes.byref(ctx)) == 1
    assert L.fl_create(1000.0, 16, 16, 0.0, ctypes.byref(ctx)) == 1
    assert L.fl_set_solid(None, 0, 0, 1) == 1


def test_splitmix64_published_vector():
    # reference output of splitmix64.c for seed 1234567
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
            16408922859458223821]
    r = Rng(1234567)
    assert [r.u64() for _ in range(5)] == want
    assert list(splitmix64(1234567, 5)) == want


def test_cavity_discs_cover_five_percent():
    nx = ny = 258
    discs = cavity_discs(nx, ny, seed=3, frac=0.05, rmin=8, rmax=32)
    solid = np.zeros((nx, ny), bool)
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    for cx, cy, r in discs:
        solid |= (ii - cx) ** 2 + (jj - cy) ** 2 <= r * r
    frac = solid[1:-1, 1:-1].mean()
    assert 0.05 <= frac < 0.2
    assert all(8 <= r <= 32 for _, _, r in discs)
