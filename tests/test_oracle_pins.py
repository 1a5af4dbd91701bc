"""The oracle passes every pin of tests/oracle_pins.py (closed forms, a typed
float32 restatement of the AC1D CPU branch, active quarter turns)."""
import pytest

import oracle_pins as P


@pytest.mark.parametrize("name", sorted(P.ALL_PINS))
def test_pin(name):
    P.ALL_PINS[name]()
