"""Each pin catches the mistake it is meant to catch.

The oracle is rebuilt with one planted mistake at a time (-DORA_SLIP=n, see the
list at the top of oracle/iac_oracle.c) into a temporary directory, and the pins
named for that mistake must fail against it.  The shipped oracle (ORA_SLIP = 0)
passes them all (tests/test_oracle_pins.py)."""
import shutil
import subprocess
from pathlib import Path

import pytest

import oracle as O
import oracle_pins as P

ORACLE_DIR = Path(O.__file__).resolve().parent


@pytest.fixture(scope="module")
def slip_libs(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    out = tmp_path_factory.mktemp("slips")
    procs = {}
    for n in P.SLIP_CAUGHT_BY:
        lib = out / f"liboracle_slip{n}.so"
        procs[n] = (lib, subprocess.Popen(
            ["gcc", "-O1", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-std=c11", "-shared",
             f"-DORA_SLIP={n}", "-o", str(lib), str(ORACLE_DIR / "iac_oracle.c"), "-lm"]))
    libs = {}
    for n, (lib, p) in procs.items():
        assert p.wait() == 0, f"slip build {n} failed"
        libs[n] = lib
    return libs


@pytest.mark.parametrize("slip", sorted(P.SLIP_CAUGHT_BY))
def test_pins_catch_planted_mistake(slip, slip_libs):
    with O.use_library(slip_libs[slip]):
        for name in P.SLIP_CAUGHT_BY[slip]:
            with pytest.raises(AssertionError):
                P.ALL_PINS[name]()
