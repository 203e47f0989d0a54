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


def manufactured(case, box, seed):
    """Diagonally dominant ADLBL3 system on `box` with a known solution xs:
    returns (ap, as_, bs, xt, xs, sl), xt = xs outside the box and 0 inside."""
    rng = np.random.default_rng(seed)
    shp = case.shape
    ap = np.ones(shp, order="F")
    as_ = np.zeros(shp + (3, 2), order="F")
    xs = np.asfortranarray(rng.uniform(-1, 1, shp))
    ib, ie, jb, je, kb, ke = box
    sl = (slice(ib - 1, ie, 2), slice(jb - 1, je, 2), slice(kb - 1, ke, 2))
    a = rng.uniform(0.1, 1.0, shp + (3, 2))
    as_[sl] = a[sl]
    ap[sl] = as_[sl].sum(axis=(-1, -2)) * rng.uniform(1.05, 1.5, ap[sl].shape)
    # b = ap*x - sum(as*x_nb)
    xsp = np.pad(xs, 2, mode="edge")
    def nb(di, dj, dk):
        return xsp[2 + di:2 + di + shp[0], 2 + dj:2 + dj + shp[1], 2 + dk:2 + dk + shp[2]]
    bs = ap * xs - (as_[..., 0, 0] * nb(-2, 0, 0) + as_[..., 0, 1] * nb(2, 0, 0) + as_[..., 1, 0] * nb(0, -2, 0)
                    + as_[..., 1, 1] * nb(0, 2, 0) + as_[..., 2, 0] * nb(0, 0, -2) + as_[..., 2, 1] * nb(0, 0, 2))
    bs = np.asfortranarray(bs)
    xt = xs.copy(order="F")
    xt[sl] = 0.0
    return ap, np.asfortranarray(as_), bs, xt, xs, sl


def wall_reference_state(cn, gold):
    """A reference run's saved state (case0041 or case1206) prepared for a
    radiation pass with its wall exchange: the records applied, the face
    values of INTP formed (the melt patches take the temperature at their
    face), the wall exchange's sbc inputs, the previous pass's mean wall
    temperature (exits of exh_type 0 use it) and the soot calibration's
    constants.  Returns (case, MinorSpecies, Radiation, GfmWallCtrl, iwall,
    fixture)."""
    import gzip
    import json
    import os

    import oracle
    from gfm_amd import radiation, sbcfile
    from gfm_amd import wall as wl
    fx = json.load(open(os.path.join(gold, "wall_fixtures.json")))["case" + cn]
    rg = gzip.open(os.path.join(gold, f"rg{cn}c.d.gz")).read()
    rs = gzip.open(os.path.join(gold, f"rs{cn}c.d.gz")).read()
    sb = os.path.join(gold, "grid", f"sbc{cn}c.dat")
    c, ms, rad = radiation.reference_state(os.path.join(gold, "grid", f"gd{cn}c.dat"), sb, rg, rs)
    oracle.intp(c)
    mp, np_, lp = c.shape
    wc, iwall = wl.inputs_from_sbc(sbcfile.read_sbc(sb), mp, np_, lp, c.arrays["ibcell"])
    wc.avg_wall_T = fx["avg_wall_T_prev"]
    if "aform_prev" in fx:
        wc.aform, wc.q_melt_req = fx["aform_prev"], fx["q_melt_req"]
    return c, ms, rad, wc, iwall, fx
