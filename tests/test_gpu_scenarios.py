"""The reference's scenario tests (fluid_test.go) on the device: the device ends bit-identical
to the float32 oracle on all eight buffers and meets every assertion the reference makes."""
import numpy as np
import pytest

from fluid_amd import Fluid
from oracle.pyoracle import OracleFluid
from scenarios import SCENARIOS

pytestmark = pytest.mark.gpu

FIELDS = ("U", "V", "newU", "newV", "p", "S", "M", "newM")


@pytest.mark.parametrize("name", list(SCENARIOS))
def test_reference_scenario_on_the_device(name):
    g, checks = SCENARIOS[name](lambda *a: Fluid(*a))
    o, _ = SCENARIOS[name](lambda *a: OracleFluid(*a))
    for k in FIELDS:
        a, b = g.get(k), o.get(k)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (name, k, int((a != b).sum()))
    assert all(checks.values()), checks
