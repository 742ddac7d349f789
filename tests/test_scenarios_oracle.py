"""The reference's scenario tests (fluid_test.go) on the float32 oracle: every assertion the
reference makes holds for the restated step."""
import pytest

from oracle.pyoracle import OracleFluid
from scenarios import SCENARIOS


@pytest.mark.parametrize("name", list(SCENARIOS))
def test_reference_scenario_on_the_oracle(name):
    _, checks = SCENARIOS[name](lambda *a: OracleFluid(*a))
    assert all(checks.values()), checks


def test_bfecc_keeps_the_peak_better():
    """TestBFECCAccuracy (fluid_test.go:1210-1268) logs the peak smoke of both methods; BFECC's
    error compensation keeps the blob's peak higher than semi-Lagrangian advection's."""
    from scenarios import bfecc_accuracy
    sl, _ = bfecc_accuracy(lambda *a: OracleFluid(*a), False)
    bf, _ = bfecc_accuracy(lambda *a: OracleFluid(*a), True)
    assert float(bf.get("M").max()) > float(sl.get("M").max())
