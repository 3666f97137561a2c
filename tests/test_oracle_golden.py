"""The CPU oracle against the reference's own golden values (tests/testthat/test-ci.R)."""
import pytest

from oracle import oracle
from tests import scenarios


@pytest.mark.parametrize("scenario", scenarios.ALL, ids=[s.__name__ for s in scenarios.ALL])
def test_oracle_matches_reference_goldens(scenario):
    scenario(oracle.edm)
