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


def test_reference_names_are_external_subroutines():
    """SPHASE, GFLOW, ADLBL3 and calc_residual are external subroutines (not
    module procedures), so the reference's implicit-interface call sites
    (C/A0main.f90:204, C/Sphase.f90:51, C/Gslv3v.f90:91-94,133,163,
    C/Gslv3g.f90:96) link to them by their plain names."""
    if not os.path.exists(EXE):
        pytest.skip("gfm_amd/fortran is not built")
    out = subprocess.run(["nm", EXE], capture_output=True, text=True, check=True).stdout
    defined = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    for name in ("sphase_", "gflow_", "adlbl3_", "calc_residual_"):
        assert name in defined, name


@pytest.mark.gpu
def test_fortran_adlbl3_reference_call_sequence(tmp_path):
    """ADLBL3(XT,IB,IE,JB,JE,KB,KE) with NELG from module state, framed by
    calc_residual as at C/Gslv3v.f90:91-94, through the Fortran host's
    external subroutines: bit-identical to the ctypes path."""
    from helpers import manufactured
    assert os.path.exists(EXE), "build gfm_amd/fortran first"
    base = cs.furnace_small(nx=16, ny=12, nz=10, seed=3)
    base.ctrl.sweep_order = 1
    mp, np_, lp = base.shape
    box = (4, mp - 2, 4, np_ - 2, 4, lp - 2)
    ap, as_, bs, xt, xs, sl = manufactured(base, box, 7)
    nelg = 4
    fin, fsys, fo = tmp_path / "in.bin", tmp_path / "sys.bin", tmp_path / "out.bin"
    write_state(base, fin)
    with open(fsys, "wb") as f:
        f.write(np.array([nelg, *box], dtype=np.int32).tobytes())
        for a in (ap, as_, bs, xt):
            f.write(np.asfortranarray(a).tobytes(order="F"))
    r = subprocess.run([EXE, "adlbl3", str(fin), str(fsys), str(fo)], capture_output=True, text=True, timeout=120,
                       cwd=str(tmp_path))
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    buf = np.fromfile(fo, dtype=np.float64)
    r0_f, r1_f = buf[0], buf[1]
    xt_f = buf[2:].reshape(base.shape, order="F")
    g = base.clone()
    xt_c = xt.copy(order="F")
    with CombustionSpace(g, initsv=False) as s:
        r0_c = s.calc_residual(ap, as_, bs, xt_c, box)
        s.adlbl3(nelg, ap, as_, bs, xt_c, box)
        r1_c = s.calc_residual(ap, as_, bs, xt_c, box)
    assert r0_f == r0_c and r1_f == r1_c
    assert r1_c < r0_c
    np.testing.assert_array_equal(xt_f, xt_c)
