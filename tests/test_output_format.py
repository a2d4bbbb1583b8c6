"""The Fortran ES16.8 edit of the AllCells writer (CPU, through the C-ABI):
every line of the reference writer's first block file (tests/golden/, from
3_6/ModSavePlot.f90 compiled with flang) is reproduced by re-editing the
values it holds, and the exponent forms Fortran uses beyond two digits."""
import gzip
import os

import pytest

import make_allcells_fixture as fx
from falcon_amd import format_es16


def _parse(field):
    """A Fortran ES field as a float (a three-digit exponent has no E)."""
    f = field.strip()
    if "E" not in f:
        k = max(f.rfind("+"), f.rfind("-"))
        f = f[:k] + "E" + f[k:]
    return float(f)


@pytest.mark.skipif(not os.path.exists(fx.FIRST), reason="fixture not generated")
def test_format_reproduces_the_reference_text():
    lines = gzip.open(fx.FIRST, "rt").read().split("\n")[:-1]
    assert len(lines) == (fx.LAYOUT["ni"] + 1) * (fx.LAYOUT["nj"] + 1) * (fx.LAYOUT["nk"] + 1)
    for ln in lines:
        assert len(ln) == 12 * 16
        assert format_es16([_parse(ln[16 * q:16 * q + 16]) for q in range(12)]) == ln


def test_format_edge_cases():
    """Right-justified width 16; three-digit exponents drop the E (the
    standard's form of ESw.d without Ee); 9.999999995 is 9.99999999499... in
    binary and must not round up."""
    assert format_es16([1.0, -1.0, 0.0]) == "  1.00000000E+00 -1.00000000E+00  0.00000000E+00"
    assert format_es16(1.5e-120) == "  1.50000000-120"
    assert format_es16(-3.3e150) == " -3.30000000+150"
    assert format_es16(9.999999995) == "  9.99999999E+00"
    assert _parse("  1.50000000-120") == 1.5e-120
