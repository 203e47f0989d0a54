"""The Fortran host (ISO_C_BINDING, gfm_amd/fortran) drives the same
library: its SPHASE must give exactly the ctypes path's result."""
import os
import subprocess

import numpy as np
import pytest

from gfm_amd import case as cs
from gfm_amd import plt, restart
from gfm_amd.solver import CombustionSpace
from gfm_amd.statefile import read_state, write_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "gfm_amd", "fortran", "gfm_comb")


def test_state_file_round_trip(tmp_path):
    c = cs.furnace_small(nx=10, ny=8, nz=8, seed=4)
    p = tmp_path / "s.bin"
    write_state(c, p)
    d = read_state(p)
    assert bytes(d.ctrl) == bytes(c.ctrl)
    for k, v in c.arrays.items():
        np.testing.assert_array_equal(d.arrays[k], v)


def test_fortran_types_match_ctypes():
    """gfm_types.f90 is generated from gfm_amd/abi.py; regenerate and compare."""
    import importlib.util
    path = os.path.join(ROOT, "gfm_amd", "fortran", "gfm_types.f90")
    before = open(path).read()
    spec = importlib.util.spec_from_file_location("gen_types", os.path.join(ROOT, "gfm_amd", "fortran", "gen_types.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    assert open(path).read() == before


@pytest.mark.gpu
def test_fortran_sphase_matches_ctypes(tmp_path):
    assert os.path.exists(EXE), "build gfm_amd/fortran first"
    base = cs.furnace_small(nx=16, ny=12, nz=10, seed=3)
    base.ctrl.sweep_order = 1
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    frg, fcv = tmp_path / "rg0001c.d", tmp_path / "convg0001c.plt"
    write_state(base, fin)
    r = subprocess.run([EXE, str(fin), str(fout), "6", "1", str(frg), str(fcv)], capture_output=True, text=True,
                       timeout=120, cwd=str(tmp_path))
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert not (tmp_path / "runend.txt").exists()
    hist_f = np.array([float(l.split()[1]) for l in r.stdout.strip().splitlines()])
    out = read_state(fout)
    g = base.clone()
    with CombustionSpace(g) as s:
        hist_c = s.sphase(6)
        s.download()
    np.testing.assert_array_equal(hist_f, hist_c)
    assert out.ctrl.itr_gas == 6
    for k in ("p", "t", "dnst"):
        np.testing.assert_array_equal(out.arrays[k], g.arrays[k])
    np.testing.assert_array_equal(out.arrays["ug"], g.arrays["ug"])
    np.testing.assert_array_equal(out.arrays["gf"], g.arrays["gf"])
    # the host's SAV3F writes the restart record the Python writer does
    assert open(frg, "rb").read() == restart.records_bytes(restart.records_from_case(out))
    # and its convg log has the reference's header and record format
    lines = open(fcv).read().split("\n")
    assert lines[:3] == plt.HEADERS["convg"]
    for i, ln in enumerate(lines[3:9]):
        assert ln[:7] == plt.fortran_i(i + 1, 7)
        assert ln[7:32] == plt.fortran_e(hist_c[i], 25, 16)
