"""The solution-file writers (alps_amd/output.py) against the reference's own
output files: every golden row, parsed and written again, is the same text
byte for byte (Fortran ES14.4E3 fields, record layout), and the row builders
put calc_eigen's outputs in the reference's column order."""
import os

import numpy as np
import pytest

from alps_amd import output

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FILES = ["test_kpar_fast.scan_kpara_1.root_1", "test_kpar_fast.eigen_kpara_1.root_1",
         "test_kpar_fast.heat_kpara_1.root_1", "test_kpar_fast.heat_mech_kpara_1.root_1"]


def fields(line, w=14):
    body = line.rstrip("\n")
    assert len(body) % w == 0, len(body)
    return [float(body[i:i + w]) for i in range(0, len(body), w)]


@pytest.mark.parametrize("name", FILES)
def test_golden_files_round_trip_byte_for_byte(name):
    text = open(os.path.join(GOLD, name)).read()
    lines = text.splitlines(keepends=True)
    assert len(lines) == 33
    out = "".join(output.row(fields(l)) for l in lines)
    assert out == text


def test_es_descriptor_cases():
    # closed forms of ES14.4E3: sign, rounding half-way cases away from the
    # binary value, exponents of one to three digits, zero
    assert output.es(0.0) == "   0.0000E+000"
    assert output.es(-0.0) == "  -0.0000E+000"
    assert output.es(1.0) == "   1.0000E+000"
    assert output.es(-2.5e-7) == "  -2.5000E-007"
    assert output.es(9.99995e10) == "   1.0000E+011"
    assert output.es(1.23456e-123) == "   1.2346E-123"
    assert output.es(123.45, 16, 6, 3) == "   1.234500E+002"
    assert output.es(float("nan")).strip() == "NaN"


def test_row_builders_follow_the_reference_column_order():
    ns = 2
    rng = np.random.default_rng(3)
    ef = rng.normal(size=3) + 1j * rng.normal(size=3)
    bf = rng.normal(size=3) + 1j * rng.normal(size=3)
    Us = rng.normal(size=(ns, 3)) + 1j * rng.normal(size=(ns, 3))  # Us(1:3, s) per row s
    ds = rng.normal(size=ns) + 1j * rng.normal(size=ns)
    w = 0.5 - 0.01j
    vals = fields(output.eigen_row(0.1, 0.2, w, ef, bf, Us, ds))
    assert len(vals) == 4 + 12 + 8 * ns
    cplx = np.array(vals[4::2]) + 1j * np.array(vals[5::2])
    # ES14.4 keeps 5 significant digits
    ref = np.concatenate([ef, bf, Us.ravel(), ds])
    assert np.allclose(cplx, ref, rtol=1e-4, atol=0)
    Ps, Psp = rng.normal(size=ns), rng.normal(size=(ns, 4))
    h = fields(output.heat_row(0.1, 0.2, w, Ps, 1.5))
    assert len(h) == 4 + ns + 1 and np.allclose(h[4:4 + ns], Ps, rtol=1e-4) and h[-1] == 1.5
    hm = fields(output.heat_mech_row(0.1, 0.2, w, Psp))
    assert len(hm) == 4 + 4 * ns and np.allclose(hm[4:], Psp.ravel(), rtol=1e-4)


def test_map_and_roots_records():
    om = np.array([[0.1 - 0.1j, 0.1 + 0.0j], [0.2 - 0.1j, 0.2 + 0.0j]])
    cal = np.array([[1e-3 + 2e-3j, -4.0 + 0j], [5e5 + 1j, 999999.0 + 0j]])
    val = np.log10(np.abs(cal))
    text = output.map_text(om, val, cal)
    lines = text.split("\n")
    assert lines[2] == "" and lines[5] == "" and len(lines[0]) == 5 * 16
    assert fields(lines[0] + "\n", 16)[:2] == [0.1, -0.1]
    r = output.roots_text([0.5 - 0.01j, 0.3 + 0j], [1e-13 + 2e-14j, 0j])
    assert r.count("\n") == 1 and r.startswith("   1") and len(r) == 4 + 5 * 14 + 1
