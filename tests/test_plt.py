"""Convergence logs (gfm_amd/plt.py) against the reference's own .plt files:
headers and record lines are reproduced byte for byte."""
import glob
import os

import pytest

from gfm_amd import case as cs
from gfm_amd import plt

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "plt")
REF = "/root/reference/GFM_Source_Package_2015-08/combustion"


def _reformat(name, text):
    """Header check and the record lines re-formatted from their values;
    returns (header_ok, records, mismatches)."""
    lines = text.split("\r\n")
    hdr = plt.HEADERS[name]
    iw, w, _ = plt.FORMATS[name]
    n = bad = 0
    for ln in lines[len(hdr):]:
        if not ln.strip():
            continue
        vals = [float(ln[iw + q * w: iw + (q + 1) * w]) for q in range((len(ln) - iw) // w)]
        n += 1
        bad += plt.format_record(name, int(ln[:iw]), vals) != ln
    return lines[:len(hdr)] == hdr, n, bad


@pytest.mark.parametrize("name,fname", [("convg", "convg0041c"), ("gresid", "gresid0041c"),
                                        ("gresid_xtra", "gresid_xtra0041c"), ("Tave", "Tave0041c")])
def test_reference_log_lines_are_reproduced(name, fname):
    """The first lines of case0041's logs: the header block and every
    record, re-formatted from the values it holds, equal the file."""
    text = open(os.path.join(GOLD, fname + ".plt"), "rb").read().decode()
    ok, n, bad = _reformat(name, text)
    assert ok and n >= 50 and bad == 0


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference data files not mounted")
def test_every_reference_log_is_reproduced():
    total = 0
    for name in plt.LOGS:
        for p in glob.glob(os.path.join(REF, "case*", f"{name}[0-9]*c.plt")):
            text = open(p, "rb").read().decode(errors="replace")
            lines = [ln for ln in text.split("\r\n")[len(plt.HEADERS[name]):] if ln.strip()]
            ok, n, bad = _reformat(name, text)
            assert bad == 0, p
            total += n
    assert total > 100000


def test_fortran_e_edge_cases():
    assert plt.fortran_e(0.0, 25, 16) == "   0.0000000000000000E+00"
    assert plt.fortran_e(-1.5, 24, 16) == " -0.1500000000000000E+01"
    assert plt.fortran_e(0.99999999999999999, 12, 4) == "  0.1000E+01"  # rounding carries into the exponent
    assert plt.fortran_e(1.0e-120, 14, 4) == "    0.1000-119"  # three-digit exponent drops the E
    assert plt.fortran_e(2.5e150, 12, 3) == "   0.250+151"
    assert plt.fortran_i(7540, 7) == "   7540"


def test_logs_written_and_appended(tmp_path):
    """A fresh run writes the headers; a restart (irstyp > 0) appends."""
    c = cs.furnace_small(nx=8, ny=6, nz=6, seed=1)
    k = c.ctrl
    k.itr_gas, k.res_mass, k.res_mass_max = 1, 0.25, 2.5
    with plt.PlotLogs(str(tmp_path), 7) as logs:
        logs.write(k)
    k.itr_gas = 2
    with plt.PlotLogs(str(tmp_path), 7, irstyp=1) as logs:
        logs.write(k)
    text = open(tmp_path / "convg0007c.plt", "rb").read().decode()
    lines = text.split("\r\n")
    assert lines[:3] == plt.HEADERS["convg"]
    assert lines[3] == "      1   0.2500000000000000E+00   0.2500000000000000E+01"
    assert lines[4].startswith("      2") and lines[5] == ""
    for n in ("gresid", "gresid_xtra", "Tave"):
        assert (tmp_path / f"{n}0007c.plt").exists()
