"""fluid_amd: MI355X-native implementation of TheFellow/fluid's pkg/fluid simulation step.

The compute path is libfluidhip.so (hand-written HIP for gfx950) behind the C ABI in
include/fluid_hip.h; `Fluid` mirrors the Go API of pkg/fluid on top of it.
"""
from .fluid import Fluid, FluidBatch, New, Rollout, ScalarField, VectorField, device_count  # noqa: F401
from ._lib import FluidError  # noqa: F401
