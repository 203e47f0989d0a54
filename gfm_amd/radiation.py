"""Radiation transport: the inputs and results of a full radiation pass
(RAD_EMIS C/Rad_emis.f90:130-229 up to QRN1: RAD0, QRNF, QREF, QRTX/Y/Z)
and of the emission-only pass (rad_qe_only C/Rad_emis.f90:411-466), run on
the device through ``CombustionSpace.rad_emis / rad_qe_only``.

``Radiation`` holds the control block (``abi.GfmRadCtrl``: the &RADIN
inputs and the pass's sums) and the host arrays of the fields on the cell
centres (reference layout, MZ = MP/2 ...):

  qe   (MZ,NZ,LZ)      emission of each open emitter (W)
  qa   (MZ,NZ,LZ)      absorption of every cell, open or boundary (W)
  qal  (MZ,NZ,LZ,NWL)  absorption per band
  akl  (MZ,NZ,LZ,NWL)  band absorptivity EMF(kappa, OPL) of open cells,
                       the emissivities of melt-surface and marked wall cells
"""
import re

import numpy as np

from .abi import NWL_MAX, GfmRadCtrl

DNST_SOOT = 1.8e3  # soot density (kg/m3), C/Var_mod.f90:137


def default_ctrl(**over):
    """The &RADIN defaults of SBC (C/Sbc.f90:316-330) with one band at
    the given wavelengths (none: 1 um)."""
    c = GfmRadCtrl()
    c.nwl, c.have_crown = 1, 0
    c.wl[0] = 1.0e-6
    c.ri = 1.0
    c.eps_w, c.eps_m, c.eps_c = 0.8, 0.95, 0.8
    c.height_to_ceiling = 0.0
    c.dnst_soot = DNST_SOOT
    for k, v in over.items():
        if k == "wl":
            if not 1 <= len(v) <= NWL_MAX:
                raise ValueError(f"1..{NWL_MAX} bands")
            c.nwl = len(v)
            for l, x in enumerate(v):
                c.wl[l] = x
        else:
            setattr(c, k, v)
    return c


def radin_from_sbc(sbc):
    """The &RADIN namelist and the band list that follows it in an sbc file
    (sbcfile.read_sbc(...).rest), as SBC reads them (C/Sbc.f90:316-355):
    returns default_ctrl keywords (wavelengths in m)."""
    text = "\n".join(sbc.rest)
    m = re.search(r"&RADIN(.*?)/", text, re.S | re.I)
    if not m:
        raise ValueError("no &RADIN namelist")
    vals = {}
    for key, val in re.findall(r"(\w+)\s*=\s*([^,\n]+)", m.group(1)):
        vals[key.lower()] = float(val.strip().rstrip(",").replace("d", "e").replace("D", "e"))
    nwl = int(vals.get("nwl", 1))
    nums = [float(x) for x in re.findall(r"[-+]?\d*\.?\d+(?:[eEdD][-+]?\d+)?", text[m.end():])]
    if len(nums) < nwl:
        raise ValueError("band list shorter than NWL")
    out = {"wl": [x * 1.0e-6 for x in nums[:nwl]]}
    for k in ("ri", "eps_w", "eps_m", "eps_c", "height_to_ceiling"):
        if k in vals:
            out[k] = vals[k]
    if "have_crown" in vals:
        out["have_crown"] = int(vals["have_crown"])
    return out


class Radiation:
    """Control block + host arrays of one case's radiation fields."""

    def __init__(self, case, **over):
        mp, np_, lp = case.shape
        self.ctrl = default_ctrl(**over)
        n = self.ctrl.nwl
        z = lambda *s: np.zeros(s, order="F")
        self.arrays = {"qe": z(mp // 2, np_ // 2, lp // 2), "qa": z(mp // 2, np_ // 2, lp // 2),
                       "qal": z(mp // 2, np_ // 2, lp // 2, n), "akl": z(mp // 2, np_ // 2, lp // 2, n)}

    def copy(self):
        o = Radiation.__new__(Radiation)
        o.ctrl = self.ctrl.copy()
        o.arrays = {k: np.array(v, order="F", copy=True) for k, v in self.arrays.items()}
        return o

    def sums(self):
        from .abi import RAD_SUMS
        return {k: getattr(self.ctrl, k) for k in RAD_SUMS}


def reference_state(gd, sbc, rg, rs, in_run=1):
    """A reference run's geometry and saved state for a radiation pass:
    gd, sbc -- paths of its grid and boundary-condition files; rg, rs --
    the bytes of its restart and minor-species records.  Returns (case,
    MinorSpecies, Radiation) with the records applied as INITSV and
    INITSVM read them and the &RADIN inputs of the sbc file.  in_run = 1
    (a later global iteration of the run, as the one that saved the
    records) takes the melt-surface cells' temperature from the record;
    0 (the first of a process) leaves it to the melt model's file."""
    from . import case as cs
    from . import gridfile, minor, restart, sbcfile
    g = gridfile.read_gd(gd)
    s = sbcfile.read_sbc(sbc)
    c = cs.from_grid(g, s.number("GMFR"))
    c.ctrl.itr_gas = 0
    restart.apply_records(c, restart.read_records(rg, *c.shape), in_run=in_run)
    # the inlet cells' temperature as SBC sets it before the records are read (C/Sbc.f90:90-105; the first
    # block holding a cell sets it); a restart reads only k and epsilon there
    ib, t = c.arrays["ibcell"], c.arrays["t"]
    taken = np.zeros(ib.shape, dtype=bool)
    for inl in s.inlets:
        i1, j1, k1, i2, j2, k2 = inl.box
        for I in range(i1, i2 + 1, 2):
            for J in range(j1, j2 + 1, 2):
                for K in range(k1, k2 + 1, 2):
                    q = (I - 1, J - 1, K - 1)
                    if ib[q] == 2 and not taken[q]:
                        taken[q] = True
                        t[q] = inl.T / c.ctrl.t0
    ms = minor.MinorSpecies(c)
    ms.load_rs(restart.read_rs(rs, *c.shape))
    return c, ms, Radiation(c, **radin_from_sbc(s))
