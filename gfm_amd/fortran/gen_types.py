"""Writes gfm_types.f90: ISO_C_BINDING mirrors of the C structs in
include/gfm_hip.h, generated from the ctypes mirror gfm_amd/abi.py so the
three stay identical.  Run: python gfm_amd/fortran/gen_types.py"""
import ctypes as ct
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from gfm_amd.abi import (GfmCtrl, GfmHostArrays, GfmMsArrays, GfmMsCtrl, GfmRadArrays, GfmRadCtrl,  # noqa: E402
                         GfmWallArrays, GfmWallCtrl)


def ftype(t):
    """(Fortran type, dimensions) of a ctypes field; C arrays of arrays map to
    Fortran arrays with the dimensions reversed (column-major)."""
    if t is ct.c_int32:
        return "integer(c_int32_t)", []
    if t is ct.c_double:
        return "real(c_double)", []
    if hasattr(t, "_length_"):
        base, dims = ftype(t._type_)
        return base, dims + [t._length_]
    if issubclass(t, ct._Pointer):
        return "type(c_ptr)", []
    raise TypeError(t)


def emit(cls, fname):
    out = [f"  type, bind(C) :: {fname}"]
    for name, t in cls._fields_:
        ty, dims = ftype(t)
        out.append(f"    {ty} :: {name}" + (f"({','.join(str(d) for d in dims)})" if dims else ""))
    out.append(f"  end type {fname}")
    return out


lines = ["! Generated by gen_types.py from gfm_amd/abi.py -- do not edit.",
         "! ISO_C_BINDING mirrors of gfm_ctrl, gfm_host_arrays, gfm_ms_ctrl, gfm_ms_arrays, gfm_rad_ctrl,",
         "! gfm_rad_arrays, gfm_wall_ctrl and gfm_wall_arrays",
         "! (include/gfm_hip.h).",
         "module gfm_types", "  use iso_c_binding", "  implicit none"]
lines += emit(GfmCtrl, "gfm_ctrl") + emit(GfmHostArrays, "gfm_host_arrays")
lines += emit(GfmMsCtrl, "gfm_ms_ctrl") + emit(GfmMsArrays, "gfm_ms_arrays")
lines += emit(GfmRadCtrl, "gfm_rad_ctrl") + emit(GfmRadArrays, "gfm_rad_arrays")
lines += emit(GfmWallCtrl, "gfm_wall_ctrl") + emit(GfmWallArrays, "gfm_wall_arrays")
lines += ["end module gfm_types", ""]
open(os.path.join(HERE, "gfm_types.f90"), "w").write("\n".join(lines))
