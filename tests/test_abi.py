"""CPU checks of the C-ABI: the library and the oracle load, export every
symbol the headers in include/ declare, and agree with the ctypes mirror on
the size of the scalar block.  No compute calls (no GPU needed)."""
import ctypes as ct
import os
import re
import subprocess

import pytest

from gfm_amd import solver
from gfm_amd.abi import (GfmCtrl, GfmHostArrays, GfmMsArrays, GfmMsCtrl, GfmRadArrays, GfmRadCtrl, GfmWallArrays,
                         GfmWallCtrl)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"\b(gfm_(?:hip|pnet)_\w+)\s*\(", txt))
    return sorted(names)


def test_library_exports_header_symbols():
    lib = solver.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", solver.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (gfm_(?:hip|pnet)_\w+)", out))
    missing = [s for s in declared() if s not in exported]
    assert not missing, missing
    for s in declared():
        getattr(lib, s)


def test_ctrl_layout():
    lib = solver.lib()
    assert lib.gfm_hip_ctrl_size() == ct.sizeof(GfmCtrl)
    assert lib.gfm_hip_abi_version() == 6
    assert ct.sizeof(GfmHostArrays) == 28 * 8
    lib.gfm_hip_ms_ctrl_size.restype = ct.c_size_t
    assert lib.gfm_hip_ms_ctrl_size() == ct.sizeof(GfmMsCtrl)
    assert ct.sizeof(GfmMsArrays) == 5 * 8
    assert lib.gfm_hip_rad_ctrl_size() == ct.sizeof(GfmRadCtrl)
    assert ct.sizeof(GfmRadArrays) == 4 * 8
    assert lib.gfm_hip_wall_ctrl_size() == ct.sizeof(GfmWallCtrl)
    assert ct.sizeof(GfmWallArrays) == 24 * 8 + 8


def test_oracle_layout():
    import oracle
    assert oracle.lib().oracle_ctrl_size() == ct.sizeof(GfmCtrl)
    assert oracle.lib().oracle_ms_ctrl_size() == ct.sizeof(GfmMsCtrl)
    oracle.lib().oracle_rad_ctrl_size.restype = ct.c_size_t
    assert oracle.lib().oracle_rad_ctrl_size() == ct.sizeof(GfmRadCtrl)
    oracle.lib().oracle_wall_ctrl_size.restype = ct.c_size_t
    assert oracle.lib().oracle_wall_ctrl_size() == ct.sizeof(GfmWallCtrl)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(solver, "_lib", None)
    monkeypatch.setattr(solver, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(solver.GfmError):
        solver.lib()


def test_noalloc_rejects_oversized_box_before_any_device_call():
    """The allocation-free calc_residual checks its box against the static
    partial-sum array (one entry per (J,K) row) before any HIP call, so the
    rejection runs without a GPU.  The pointers are never dereferenced."""
    lib = solver.lib()
    f = lib.gfm_hip_calc_residual_noalloc
    f.argtypes = [ct.c_int] * 4 + [ct.c_void_p] * 5 + [ct.c_int] * 6 + [ct.c_void_p] * 2
    f.restype = ct.c_int
    fake = ct.c_void_p(16)
    n = 600  # 297 x 297 rows > 65536
    rc = f(8, n, n, 1, fake, fake, fake, fake, fake, 3, 5, 3, n - 3, 3, n - 3, fake, None)
    assert rc == 1
    assert b"65536" in lib.gfm_hip_last_error()
    rc = f(8, 20, 20, 1, fake, fake, fake, fake, fake, 3, 6, 3, 17, 3, 17, fake, None)  # odd box extent
    assert rc == 1


def test_minor_species_entry_points_reject_bad_controls_before_any_device_call():
    """The gfm_hip_ms_* entry points check the control block first: a null
    context or a relaxation factor outside (0, 1] is rejected without a
    GPU."""
    lib = solver.lib()
    from gfm_amd import case as cs, minor
    c = cs.furnace_small(nx=8, ny=6, nz=6, seed=1)
    m = minor.default_ctrl(c)
    for f in (lib.gfm_hip_gslv3m, lib.gfm_hip_msflow):
        f.argtypes = [ct.c_void_p, ct.POINTER(GfmMsCtrl)]
        assert f(None, ct.byref(m)) == 1
    bad = m.copy()
    bad.rfm[3] = 0.0
    lib.gfm_hip_initsvm.argtypes = [ct.c_void_p, ct.POINTER(GfmMsCtrl), ct.c_int]
    assert lib.gfm_hip_initsvm(ct.c_void_p(1), ct.byref(bad), 0) == 1
