"""The reference's scenario tests (fluid_test.go) on the float32 oracle: every assertion the
reference makes holds for the restated step."""
import pytest

from oracle.pyoracle import OracleFluid
from scenarios import SCENARIOS


@pytest.mark.parametrize("name", list(SCENARIOS))
def test_reference_scenario_on_the_oracle(name):
    _, checks = SCENARIOS[name](lambda *a: OracleFluid(*a))
    assert all(checks.values()), checks
