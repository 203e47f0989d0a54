"""Minor species and soot: the state and the inputs of the reference's
sub-species solve (MSFLOW C/Msflow.f90:10, GSLV3M C/Gslv3m.f90:8, SOURCEM
C/Sourcem.f90:8, INITSVM C/Initsv.f90:329), run on the device through
``CombustionSpace.initsvm / gslv3m / msflow``.

``MinorSpecies`` holds the control block (``abi.GfmMsCtrl``) and the host
arrays of the state in the reference layout:

  gfm            (MP,NP,LP,7)  CH4, O2, CO, H2O, CO2, N2, NO mass fractions
  smf            (MP,NP,LP)    soot mass fraction
  bs_soot_prev   (MP,NP,LP)    soot formation source of the last SOURCEM(11)
  sfp_soot_prev  (MP,NP,LP)    soot oxidation source of the last SOURCEM(11)
  rrms           (MP/2,NP/2,LP/2,5)  reaction rates (kmol/s), 1, 2, 4 formed
"""
import numpy as np

from .abi import GfmMsCtrl

RU = 8.3143e3  # universal gas constant (J/kmol/K), C/Var_mod.f90:22


def vol_tot(case):
    """VOL_TOT of GRID3 (C/Grid3.f90:286-294): the volume (m3) of the cells
    with code 0, summed in the reference's loop order (I, J, then K)."""
    A = case.arrays
    ib = A["ibcell"][1::2, 1::2, 1::2]
    dx, dr, rs, dz = A["dx"][1::2], A["dr"][1::2], A["rs"][1::2], A["dz"][1::2]
    v = ((dx[:, None, None] * dr[None, :, None]) * rs[None, :, None]) * dz[None, None, :]
    v = np.where(ib == 0, v * case.ctrl.vol0, 0.0)
    return float(np.cumsum(v.ravel(order="C"))[-1])  # C order: K fastest, I slowest


def default_ctrl(case, **over):
    """The reference defaults: SETUP (C/Setup.f90:161-185) and the shipped
    relaxation file (Release_files/base_relaxfactorc.txt: minor species 0.5,
    soot 0.2, soot sources 0.3, soot cap 2e-3, cut-off temperatures 0,
    4 sweeps)."""
    m = GfmMsCtrl()
    m.maxms, m.react, m.ntimesm, m.itr_ms, m.igs0 = 300, 1, 4, 0, 0
    m.bmcon, m.ssigma = 1.0e-7, 0.7
    for l in range(7):
        m.rfm[l] = 0.5
    m.rfm[7] = 0.2
    m.aform, m.aoxid = 400.0, 26.0
    m.esf_r = m.eso_r = 5.0e7 / RU
    m.t0_sf = m.t0_so = 0.0
    m.rf_soot_form_source = m.rf_soot_oxid_source = 0.3
    m.soot_cap, m.dnst_soot = 2.0e-3, 1.8e3
    m.vol_tot = vol_tot(case)
    for k, v in over.items():
        if k == "rfm":
            for l, x in enumerate(v):
                m.rfm[l] = x
        else:
            setattr(m, k, v)
    return m


class MinorSpecies:
    """Control block + host arrays of one case's minor-species state."""

    def __init__(self, case, **over):
        mp, np_, lp = case.shape
        self.ctrl = default_ctrl(case, **over)
        z = lambda *s: np.zeros(s, order="F")
        self.arrays = {"gfm": z(mp, np_, lp, 7), "smf": z(mp, np_, lp), "bs_soot_prev": z(mp, np_, lp),
                       "sfp_soot_prev": z(mp, np_, lp), "rrms": z(mp // 2, np_ // 2, lp // 2, 5)}

    def load_rs(self, rec):
        """Put an rs####c.d record (restart.read_rs) into the arrays, as
        INITSVM's reads do (C/Initsv.f90:352-372); the inlet override and
        extrm are INITSVM's restart branch (initsvm with restart=1)."""
        from .restart import _cell_index
        mp, np_, lp = self.arrays["smf"].shape
        if (rec.mp, rec.np, rec.lp) != (mp, np_, lp):
            raise ValueError("restart record and case lattices differ")
        I, J, K, _, _, _ = _cell_index(mp, np_, lp)
        self.arrays["gfm"][I, J, K, :] = rec.gfm
        self.arrays["smf"][I, J, K] = rec.smf
        self.arrays["bs_soot_prev"][...] = rec.bs_soot_prev
        self.arrays["sfp_soot_prev"][...] = rec.sfp_soot_prev
        self.arrays["rrms"][...] = rec.rrms
        self.ctrl.itr_ms = int(rec.itr_ms)

    def copy(self):
        o = MinorSpecies.__new__(MinorSpecies)
        o.ctrl = self.ctrl.copy()
        o.arrays = {k: np.array(v, order="F", copy=True) for k, v in self.arrays.items()}
        return o
