"""Comparison helpers shared by the parity tests."""
import numpy as np

from gfm_amd.case import IYF, IYO2, IH, IK, IEPS, IYCO2, IYN2, IYH2O


def open_centres(case):
    ib = case.arrays["ibcell"]
    m = np.zeros(ib.shape, dtype=bool)
    m[1::2, 1::2, 1::2] = ib[1::2, 1::2, 1::2] <= 0
    return m


def open_faces(case, iu):
    ib = case.arrays["ibcell"]
    m = np.zeros(ib.shape, dtype=bool)
    sl = [slice(1, None, 2)] * 3
    sl[iu - 1] = slice(2, None, 2)
    sl = tuple(sl)
    m[sl] = ib[sl] <= 0
    return m


def field_errors(a, b):
    """Max relative error per field between two cases, over the positions
    each field is solved on (open cell centres; velocities on their own
    open faces).  Scale: max |value| of the field over those positions."""
    out = {}
    mc = open_centres(a)
    for name, arr_a, arr_b in [("p", a.arrays["p"], b.arrays["p"]), ("t", a.arrays["t"], b.arrays["t"]),
                               ("dnst", a.arrays["dnst"], b.arrays["dnst"])]:
        out[name] = _rel(arr_a[mc], arr_b[mc])
    names = {IYF: "yf", IYO2: "yo2", IH: "h", IK: "k", IEPS: "eps", IYCO2: "yco2", IYN2: "yn2", IYH2O: "yh2o"}
    for l, nm in names.items():
        out[nm] = _rel(a.arrays["gf"][..., l - 1][mc], b.arrays["gf"][..., l - 1][mc])
    for iu in (1, 2, 3):
        mf = open_faces(a, iu)
        out["uvw"[iu - 1]] = _rel(a.arrays["ug"][..., iu - 1][mf], b.arrays["ug"][..., iu - 1][mf])
    return out


def _rel(x, y):
    scale = max(np.max(np.abs(x)), np.max(np.abs(y)), 1e-300)
    return float(np.max(np.abs(x - y)) / scale)
