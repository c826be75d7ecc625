"""The oracle's functions that no reference-held output covers, pinned by
closed forms, textbook formulas and brute-force restatements
(tests/oracle_pins.py describes each)."""
import pytest

import oracle_pins


@pytest.mark.parametrize("name", sorted(oracle_pins.CHECKS))
def test_oracle_pin(oracle, name):
    oracle_pins.CHECKS[name](oracle)
