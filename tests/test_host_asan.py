"""The host runtime under AddressSanitizer: libfluidhip_asan.so (fluid_amd/build.py --asan; the
same kernels, host code built with -fsanitize=address) runs in a child process with the ASan
runtime preloaded. On the CPU: loading, the exported symbols and the argument and no-device
error paths. On the GPU: tests/asan_exercise.py, which drives contexts, edits, every solver
path, views, particles, staged copies, batches, the rollout and the tuners. A heap error in
host code aborts the child with an AddressSanitizer report, which fails the test."""
import os
import subprocess
import sys

import pytest

from fluid_amd import build as fb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_sanitized(code_or_script, timeout):
    if not os.path.exists(fb.ASAN_LIB):
        pytest.fail("libfluidhip_asan.so is not built (python fluid_amd/build.py --asan)")
    env = dict(os.environ)
    env["LD_PRELOAD"] = fb.asan_runtime()
    env["FLUID_AMD_LIB"] = fb.ASAN_LIB
    # Python's own allocations at exit are not the library's: leak checking off; stop at the
    # first error
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:abort_on_error=0"
    args = [sys.executable, "-u"] + (["-c", code_or_script] if "\n" in code_or_script else [code_or_script])
    r = subprocess.run(args, cwd=ROOT, env=env, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0 or "AddressSanitizer" in r.stderr:
        raise AssertionError("sanitized run failed (rc %d):\n%s\n%s" % (r.returncode, r.stdout[-2000:], r.stderr[-6000:]))
    return r.stdout


CPU_CHECKS = """
import ctypes
from fluid_amd import _lib
L = _lib.load()
assert _lib.LIB_PATH.endswith("libfluidhip_asan.so")
for n in _lib.SIGNATURES:
    assert hasattr(L, n), n
n = ctypes.c_int(-1)
assert L.fl_device_count(ctypes.byref(n)) == 0
ctx = ctypes.c_void_p()
assert L.fl_create(1000.0, 0, 16, 0.01, ctypes.byref(ctx)) == 1
assert L.fl_create(1000.0, 16, 16, 0.0, ctypes.byref(ctx)) == 1
assert L.fl_set_solid(None, 0, 0, 1) == 1
for _ in range(100):
    L.fl_last_error()
if n.value == 0:
    assert L.fl_create(1000.0, 16, 16, 0.01, ctypes.byref(ctx)) == 4 and not ctx.value
    assert b"no HIP device" in L.fl_last_error()
print("asan cpu ok")
"""


def test_host_asan_error_paths():
    assert "asan cpu ok" in run_sanitized(CPU_CHECKS, timeout=120)


@pytest.mark.gpu
def test_host_asan_exercise():
    print("host ASan: tests/asan_exercise.py against libfluidhip_asan.so", flush=True)
    out = run_sanitized(os.path.join(ROOT, "tests", "asan_exercise.py"), timeout=280)
    assert "asan exercise ok" in out


OVERFLOW = """
import ctypes
import numpy as np
from fluid_amd import Fluid
f = Fluid(1000.0, 1100, 600, 0.01)
n = f.NumX * f.NumY
short = np.zeros(n - 1000, np.float32)
# the library is told the full field size: the host copy of its last staged chunk (no device
# copy is in flight by then) writes 4000 bytes past the buffer
f._L.fl_download(f._ctx, 0, short.ctypes.data, n)
print("no report")
"""


@pytest.mark.gpu
def test_host_asan_reports_an_overflowing_download():
    """The sanitizer is live: a download into a buffer 4000 bytes shorter than it claims (a
    2.6 MB field, copied by the library's host threads from pinned staging) stops with a
    heap-buffer-overflow report from inside the library, on the host, before the write."""
    print("host ASan: overflowing download, expected report", flush=True)
    with pytest.raises(AssertionError, match="heap-buffer-overflow"):
        run_sanitized(OVERFLOW, timeout=120)
