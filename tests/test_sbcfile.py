"""Boundary-condition files sbc####c.dat (gfm_amd/sbcfile.py): the inlet
blocks and the inflow SBC forms from them (C/Sbc.f90:19-170) against what
the reference's runs logged (C/Utility.f90:101-105,131: mass inflow, fuel
heat value, enthalpy inflow -- the last pins T_to_enth at the inlet
states), and the case builder's own inflow against the same function."""
import glob
import json
import os
from types import SimpleNamespace

import pytest

from gfm_amd import case as cs
from gfm_amd import gridfile as gfile
from gfm_amd import sbcfile as sf

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
REF = "/root/reference/GFM_Source_Package_2015-08/combustion"


def test_reference_sbc_is_parsed():
    s = sf.read_sbc(os.path.join(GOLD, "grid", "sbc0041c.dat"))
    assert s.number("GMFR") == 0.9242821 and s.number("Q0") == 5e7
    assert s.p_ref_cell == (6, 28, 10) and s.p_ref == 0.0
    assert len(s.inlets) == 3
    a = s.inlets[2]
    assert a.box == (6, 24, 16, 6, 42, 26) and a.iu == 1 and a.T == 1473.0 and a.yo2 == 0.2329175
    assert a.uvw == (7.95023505148322, -0.835603395376252, -2.8936488014473)
    assert s.rest[0].startswith("Exhausts:")


@pytest.mark.parametrize("cn", ["0041", "1206"])
def test_inflow_rates_match_the_logged_ones(cn):
    want = json.load(open(os.path.join(GOLD, "grid_fixtures.json")))[f"gd{cn}c.dat"]
    g = gfile.read_gd(os.path.join(GOLD, "grid", f"gd{cn}c.dat"))
    s = sf.read_sbc(os.path.join(GOLD, "grid", f"sbc{cn}c.dat"))
    got = sf.inflow_rates(g, s)
    for k in ("fr_f", "fr_ox", "fr_n2", "heat_in"):
        assert got[k] == pytest.approx(want[k], rel=1e-13), k
    if cn == "0041":  # the enthalpy inflow (case 1206's log does not match its own inlets)
        assert got["q_h_in"] == pytest.approx(want["q_h_in"], rel=1e-13)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference data files not mounted")
def test_every_reference_case_gives_its_logged_inflow():
    n = nq = 0
    for sm in glob.glob(os.path.join(REF, "case*", "summary*c.txt")):
        d = os.path.dirname(sm)
        gd, sb = glob.glob(os.path.join(d, "gd*c.dat")), glob.glob(os.path.join(d, "sbc*c.dat"))
        if not gd or not sb:
            continue
        logged = {}
        for ln in open(sm).read().splitlines():
            if "Fuel in flow rate" in ln:
                logged["f"] = float(ln.split(":")[1].split()[0])
            if "Total in flow rate" in ln:
                logged["t"] = float(ln.split(":")[1].split()[0])
            if "Energy in with flow" in ln and "q" not in logged:
                logged["q"] = float(ln.split(":")[1].split()[0])
        r = sf.inflow_rates(gfile.read_gd(gd[0]), sf.read_sbc(sb[0]))
        assert r["fr_f"] == pytest.approx(logged["f"], rel=1e-13), d
        assert r["fr_f"] + r["fr_ox"] + r["fr_n2"] == pytest.approx(logged["t"], rel=1e-13), d
        # enthalpy inflow: reproduced for every run of the current code; the
        # runs kept as "oldrun"/"bak" log 2 % less, and cases 1206/1207 log
        # 3.8e5 W for 13 kg/s of mostly 1400 K air, which their inlet files do
        # not give whatever the air temperature
        if not ("oldrun" in d or "bak" in d or "case120" in d):
            assert r["q_h_in"] == pytest.approx(logged["q"], rel=1e-13), d
            nq += 1
        n += 1
    assert n >= 15 and nq >= 10


@pytest.mark.parametrize("make", [lambda: cs.furnace_small(nx=16, ny=12, nz=10, seed=3),
                                  lambda: cs.furnace_large(nx=48, ny=24, nz=16, seed=2),
                                  lambda: cs.furnace_small(fuel_u=33.0, air_t=1000.0)])
def test_case_builder_inflow_is_sbc_inflow(make):
    """The furnaces built here set their inlets with the same sum: the
    builder's fuel, O2 and N2 inflow equal the pinned function's."""
    c = make()
    a = c.arrays
    g = SimpleNamespace(ibcell=a["ibcell"], dx=a["dx"], dr=a["dr"], dz=a["dz"])
    s = sf.SbcFile({"GMFR": repr(c.ctrl.gmfr), "Q0": repr(c.spec.q0)}, (c.ctrl.i_ref, c.ctrl.j_ref, c.ctrl.k_ref),
                   0.0, c.spec.inlets)
    r = sf.inflow_rates(g, s)
    assert (r["fr_f"], r["fr_ox"], r["fr_n2"]) == pytest.approx((c.meta["fr_f"], c.meta["fr_ox"], c.meta["fr_n2"]),
                                                                rel=1e-15)
