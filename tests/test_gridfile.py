"""Grid files gd####c.dat (gfm_amd/gridfile.py) against GRID3's own record
of the reference's cases: the open-cell count and melt-surface area each
run logged in its summary file (C/Grid3.f90:257-288, C/Utility.f90:99-100)."""
import glob
import json
import os

import numpy as np
import pytest

from gfm_amd import case as cs
from gfm_amd import gridfile as gfile

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
REF = "/root/reference/GFM_Source_Package_2015-08/combustion"


@pytest.mark.parametrize("name", ["gd0041c.dat", "gd1206c.dat"])
def test_reference_grid_gives_the_logged_cells_and_melt_area(name):
    want = json.load(open(os.path.join(GOLD, "grid_fixtures.json")))[name]
    g = gfile.read_gd(os.path.join(GOLD, "grid", name))
    assert g.ncells == want["ncells"]
    assert g.area_melt_surf == pytest.approx(want["area_melt_surf"], rel=1e-13)
    mp, np_, lp = g.shape
    assert g.x[mp - 1] == g.x[mp - 2] and g.dx[0] == g.dx[1] == 0.0 and np.all(g.dx[2:mp - 1] > 0)
    # propagation: every lattice neighbour of a wall centre is a wall
    w = np.argwhere(g.centre_codes == 1)[0] * 2 + 1
    assert np.all(g.ibcell[w[0] - 1:w[0] + 2, w[1] - 1:w[1] + 2, w[2] - 1:w[2] + 2] == 1)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference data files not mounted")
def test_every_reference_grid_gives_its_logged_cells():
    n = 0
    for sm in glob.glob(os.path.join(REF, "case*", "summary*c.txt")):
        gd = glob.glob(os.path.join(os.path.dirname(sm), "gd*c.dat"))
        if not gd:
            continue
        lines = open(sm).read().splitlines()
        g = gfile.read_gd(gd[0])
        assert g.ncells == int(lines[2].split(":")[1]), gd[0]
        assert g.area_melt_surf == pytest.approx(float(lines[3].split(":")[1].split()[0]), rel=1e-13), gd[0]
        n += 1
    assert n >= 15


@pytest.mark.parametrize("make", [lambda: cs.furnace_small(nx=16, ny=12, nz=10, seed=3),
                                  lambda: cs.furnace_large(nx=48, ny=24, nz=16, seed=2)])
def test_written_grid_reads_back(tmp_path, make):
    """A furnace built here, written as a gd file and read as GRID3 reads
    it: the same cell codes everywhere on the lattice and the same
    coordinates up to GRID3's cumulative step (the stretched furnace lists
    every face)."""
    c = make()
    p = tmp_path / "gd0001c.dat"
    gfile.write_gd(str(p), c)
    g = gfile.read_gd(str(p))
    np.testing.assert_array_equal(g.ibcell, c.arrays["ibcell"])
    assert g.ncells == c.ctrl.ncells
    for k in ("x", "dx", "r", "dr", "z", "dz"):
        np.testing.assert_allclose(getattr(g, k), c.arrays[k], rtol=0, atol=1e-14 * np.max(np.abs(c.arrays[k])))


def test_malformed_grid_is_rejected():
    text = open(os.path.join(GOLD, "grid", "gd0041c.dat")).read()
    lines = text.split("\n")
    bad = list(lines)
    i = next(q for q, ln in enumerate(lines) if ln.startswith("X grid points"))
    bad[i + 2] = bad[i + 2].replace("0.", "-9.", 1)  # coordinates no longer increasing
    with pytest.raises(ValueError):
        gfile.read_gd("\n".join(bad))
    bad = list(lines)
    j = next(q for q, ln in enumerate(lines) if ln.startswith("Computational")) + 2
    bad[j] = bad[j][:-3]  # a cell-map row short of digits
    with pytest.raises(ValueError):
        gfile.read_gd("\n".join(bad))


def test_half_blocked_cells_round_trip_with_their_factors(tmp_path):
    """Cells of code -1 and -2 are written as digits 9 and 8 and the &GRD
    factors BV_F, BA_F as a namelist; GRID3's reading rules give them back
    (C/Grid3.f90:243-256), as open cells for NCELLS."""
    hb = [((6, 8), (3, 5), (2, 4), -1), ((15, 15), (7, 9), (6, 7), -2)]
    bv, ba = (0.5, 0.7), (0.6, 0.8, 0.9, 0.75, 0.65, 0.85)
    c = cs.furnace_small(nx=16, ny=12, nz=10, seed=3, half_blocked=hb, bv_f=bv, ba_f=ba)
    p = tmp_path / "gd0001c.dat"
    gfile.write_gd(p, c)
    text = p.read_text()
    assert "9" in text.split("Computational and blocked cells")[1].split("&GRD")[0]
    g = gfile.read_gd(str(p))
    np.testing.assert_array_equal(g.ibcell, c.arrays["ibcell"])
    assert g.ncells == c.ctrl.ncells
    assert gfile.grd_factors(g.grd) == (bv, ba)
    # element forms of the namelist
    assert gfile.grd_factors({"BA_F(2,2)": [0.5], "BV_F(2)": [0.25]}) == ((1.0, 0.25), (1.0,) * 4 + (0.5, 1.0))
