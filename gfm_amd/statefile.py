"""Stream-file exchange of a case state with the Fortran host
(gfm_amd/fortran/gfm_comb.f90): magic 'GFMSTATE', version, sizeof(ctrl),
the ctrl block, then the arrays in reference (column-major) layout."""
import ctypes as ct

import numpy as np

from .abi import GfmCtrl, FIELD3, GRID1, KIN
from .case import Case

ORDER3 = ("p", "t", "dnst", "theta", "tmu", "gcp", "glam", "gama", "gte", "crate", "hrate", "bs_hprev")


def write_state(case, path):
    with open(path, "wb") as f:
        f.write(b"GFMSTATE")
        f.write(np.array([1, ct.sizeof(GfmCtrl)], dtype=np.int32).tobytes())
        f.write(bytes(case.ctrl))
        for k in GRID1:
            f.write(np.ascontiguousarray(case.arrays[k], dtype=np.float64).tobytes())
        f.write(case.arrays["ibcell"].tobytes(order="F"))
        for k in ORDER3:
            f.write(case.arrays[k].tobytes(order="F"))
        f.write(case.arrays["ug"].tobytes(order="F"))
        f.write(case.arrays["gf"].tobytes(order="F"))
        for k in KIN:
            f.write(np.ascontiguousarray(case.arrays[k], dtype=np.float64).tobytes())


def read_state(path, spec=None):
    buf = open(path, "rb").read()
    assert buf[:8] == b"GFMSTATE", "not a state file"
    ver, csize = np.frombuffer(buf[8:16], dtype=np.int32)
    assert csize == ct.sizeof(GfmCtrl)
    ctrl = GfmCtrl.from_buffer_copy(buf[16:16 + csize])
    off = 16 + csize
    mp, np_, lp, nr0 = ctrl.mp, ctrl.np, ctrl.lp, ctrl.nr0
    n = mp * np_ * lp

    def take(count, dtype):
        nonlocal off
        a = np.frombuffer(buf, dtype=dtype, count=count, offset=off).copy()
        off += count * np.dtype(dtype).itemsize
        return a

    arr = {}
    for k, m in zip(GRID1, (mp, mp, np_, np_, np_, lp, lp)):
        arr[k] = take(m, np.float64)
    arr["ibcell"] = take(n, np.int16).reshape((mp, np_, lp), order="F")
    for k in ORDER3:
        arr[k] = take(n, np.float64).reshape((mp, np_, lp), order="F")
    arr["ug"] = take(3 * n, np.float64).reshape((mp, np_, lp, 3), order="F")
    arr["gf"] = take(9 * n, np.float64).reshape((mp, np_, lp, 9), order="F")
    for k in KIN:
        arr[k] = take(nr0, np.float64)
    return Case(ctrl, arr, spec)
