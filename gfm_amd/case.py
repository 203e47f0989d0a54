"""Synthetic furnace cases in the reference's data model.

The reference builds the state that SPHASE consumes in four host routines:
GRID3 (grid and cell codes, C/Grid3.f90:8-300), SETUP (reference scales
and nondimensional groups, C/Setup.f90:223-731), SBC (inlet cells,
C/Sbc.f90:19-394) and INITSV/gen_initial_guess (initial fields,
C/Initsv.f90:30-209).  This module restates them for the synthetic cases
the benchmarks run on, producing the doubled-lattice host arrays and the
scalar block of the C-ABI (gfm_amd.abi).  The tail of INITSV (INTPV,
EXTR(1), INTP) is part of the solver path and runs in the library
(``gfm_hip_initsv_tail``).

C/ = combustion/Code/ of the reference package.
"""
from dataclasses import dataclass, field
import math
import os

import numpy as np

from .abi import GfmCtrl, FIELD3

# ---- constants of GBL_VAR, C/Var_mod.f90 ----
RU = 8.3143e3
WF, WO2, WCO2, WH2O, WN2 = 16.0426, 31.9988, 44.0098, 18.0152, 28.0134
WT0 = 1.0 / (0.5 / WF + 0.0 / WCO2 + 0.0 / WH2O + 0.49 / WO2 + 0.01 / WN2)
T0 = 500.0
TL0 = 298.15
CP0 = 1.0057e3
H_0 = CP0 * (T0 - TL0)
CMU = 9.0e-2
CTEA, CTDA = 5.0e-2, 1.0
PG0 = 1.01325e5
GMU_REF = 2.6e-5
CPF1 = 2.156e3 / 1.0057e3
IYF, IYO2, IH, IK, IEPS, IYCO2, IYN2, IYH2O = 1, 2, 3, 4, 5, 6, 7, 9

# NIST Shomate coefficients (C/Setup.f90:103-141), species fuel,O2,CO2,N2,H2O
A_N = np.array([[-0.703029, 85.81217], [29.659, 29.659], [24.99735, 58.16639], [26.092, 26.092], [30.092, 41.96426]])
B_N = np.array([[108.4773, 11.26467], [6.137261, 6.137261], [55.18696, 2.720074], [8.218801, 8.218801], [6.832514, 8.622053]])
C_N = np.array([[-42.52157, -2.114146], [-1.186521, -1.186521], [-33.69137, -0.492289], [-1.976141, -1.976141], [6.793435, -1.49978]])
D_N = np.array([[5.862788, 0.13819], [0.09578, 0.09578], [7.948387, 0.038844], [0.159274, 0.159274], [-2.53448, 0.098119]])
E_N = np.array([[0.678565, -26.42221], [-0.219663, -0.219663], [-0.136638, -6.447293], [0.044434, 0.044434], [0.082139, -11.15764]])
F_N = np.array([[-76.84376, -153.5327], [-9.861391, -9.861391], [-403.6075, -425.9186], [-7.98923, -7.98923], [-250.881, -272.1797]])
H_N = np.array([[-74.8731, -74.8731], [0.0, 0.0], [-393.5224, -393.5224], [0.0, 0.0], [-241.8264, -241.8264]])
CUT_T = np.array([1300.0, 6000.0, 1200.0, 6000.0, 1700.0])
CHEM = (IYF, IYO2, IYCO2, IYN2, IYH2O)

# shipped relaxation/sweep settings, combustion/relaxfactorc.txt read by C/Setup.f90:438-515
RFL = {1: 0.6, 2: 0.6, 3: 0.2, 4: 0.6, 5: 0.6, 6: 0.6, 7: 0.6}
RFL_GENERAL = 0.6
RF_GENERAL = 0.7
RF = {1: 0.3, 2: 0.3, 3: 0.3, 18: 0.6, 21: 0.6}
RF_HSOURCE = 0.3
NTIMESG = [2, 2, 2, 6, 5, 3, 3, 8, 2, 2, 3, 3] + [2] * 8

KINETIC_FILE = os.path.join(os.path.dirname(__file__), "data", "kinetic.d")


def t_to_enth(y5, tn):
    """T_to_enth, C/Prop.f90:525-565 (vectorised), including the
    non-accumulated f_mix/h_mix construction of the reference."""
    y5 = np.asarray(y5, dtype=np.float64)
    tn = np.asarray(tn, dtype=np.float64)
    w_mix = 1.0 / (y5[0] / WF + y5[2] / WCO2 + y5[4] / WH2O + y5[1] / WO2 + y5[3] / WN2)
    real_t = tn * T0
    a = np.zeros_like(tn); b = np.zeros_like(tn); c = np.zeros_like(tn); d = np.zeros_like(tn); e = np.zeros_like(tn)
    f = np.zeros_like(tn); h = np.zeros_like(tn)
    for n in range(5):
        s = np.where(real_t < CUT_T[n], 0, 1)
        a = a + y5[n] * A_N[n][s]
        b = b + y5[n] * B_N[n][s]
        c = c + y5[n] * C_N[n][s]
        d = d + y5[n] * D_N[n][s]
        e = e + y5[n] * E_N[n][s]
        f = d + y5[n] * F_N[n][s]
        h = e + y5[n] * H_N[n][s]
    t = real_t / 1000.0
    ef = t * (a + t * ((b / 2.0) + t * ((c / 3.0) + t * (d / 4.0)))) - (e / t) + f - h
    return ef * 1.0e6 / (w_mix * H_0)


def mixture_w(y5):
    y5 = np.asarray(y5, dtype=np.float64)
    return 1.0 / (y5[0] / WF + y5[2] / WCO2 + y5[4] / WH2O + y5[1] / WO2 + y5[3] / WN2)


@dataclass
class Inlet:
    """One inlet block of the sbc file (C/Sbc.f90:67-70): lattice index box
    (I1..I2, J1..J2, K1..K2), normal direction IU and the SI state."""
    box: tuple
    iu: int
    T: float
    yf: float
    yo2: float
    yn2: float
    yco2: float = 0.0
    yh2o: float = 0.0
    uvw: tuple = (0.0, 0.0, 0.0)


@dataclass
class CaseSpec:
    name: str
    faces_x: np.ndarray
    faces_y: np.ndarray
    faces_z: np.ndarray
    inlets: list
    exit_face: str = "x+"
    exit_yz: tuple = None        # (y0, y1, z0, z1) port on the x+ end wall; None -> full face
    t_init: float = 1500.0
    k_init: float = None        # SI m2/s2; None -> TKE() from velocity
    eps_init: float = None      # SI m2/s3
    noise: float = 0.0          # multiplicative uniform noise on T, k, eps
    seed: int = 0
    q0: float = 5.0e7           # heat of combustion, J/kg fuel
    melt_bottom_T: object = None  # callable x->T (K) for an IBCELL=4 bottom, or None
    t_wall: float = None        # isothermal wall temperature (K), None -> adiabatic (reference)
    gray_kappa: float = 0.0     # gray absorption coefficient (1/m) of the radiation source term; 0 -> off
    bgcon: float = 1.0e-5
    maxgi: int = 2000
    # half-blocked cells (grid-file digits 9 and 8, C/Grid3.f90:243-256):
    # list of ((c0, c1), (d0, d1), (e0, e1), code) with 0-based inclusive
    # cell ranges and code -1 or -2; the &GRD factors BV_F(1:2) and
    # BA_F(1:3,1:2) (column-major) of their volumes and face areas
    half_blocked: list = field(default_factory=list)
    bv_f: tuple = (1.0, 1.0)
    ba_f: tuple = (1.0,) * 6
    meta: dict = field(default_factory=dict)


def lattice_coord(faces):
    """Doubled-lattice coordinate array from cell-face coordinates, as GRID3
    builds it from the gd file (C/Grid3.f90:79-109): X(1..3)=x0, odd
    indices are faces, even indices are centres, X(MP)=X(MP-1)."""
    faces = np.asarray(faces, dtype=np.float64)
    n = len(faces) - 1
    mp = 2 * n + 4
    x = np.zeros(mp + 1)  # 1-based
    x[1] = x[2] = x[3] = faces[0]
    for c in range(1, n + 2):
        x[2 * c + 1] = faces[c - 1]
    for c in range(2, n + 2):
        x[2 * c] = 0.5 * (x[2 * c - 1] + x[2 * c + 1])
    x[mp] = x[mp - 1]
    dx = np.zeros(mp + 1)
    for i in range(2, mp):
        dx[i] = x[i + 1] - x[i - 1]
    return x[1:].copy(), dx[1:].copy()


def propagate_ibcell(ib):
    """26-neighbour propagation of cell codes onto faces, edges and corners,
    walls sticky, in the reference's loop order (C/Grid3.f90:259-275).
    ``ib`` is 0-based (MP,NP,LP) with centre codes set; returns NCELLS."""
    mp, np_, lp = ib.shape
    cen = ib[1::2, 1::2, 1::2]
    ncells = int(np.count_nonzero(cen <= 0))
    idx = np.argwhere(cen >= 1)  # lexicographic (I,J,K) = Fortran I-outer order
    for ci, cj, ck in idx:
        i, j, k = 2 * ci + 1, 2 * cj + 1, 2 * ck + 1  # 0-based of I=2ci+2
        code = ib[i, j, k]
        for i1 in (i - 1, i, i + 1):
            if i1 >= mp:
                continue
            for j1 in (j - 1, j, j + 1):
                if j1 >= np_:
                    continue
                for k1 in (k - 1, k, k + 1):
                    if k1 >= lp:
                        continue
                    if ib[i1, j1, k1] != 1:
                        ib[i1, j1, k1] = code
    return ncells


def load_kinetic(nr0=100):
    """kinetic.d tables with the scalings of C/Setup.f90:858-888."""
    rows = []
    with open(KINETIC_FILE) as fh:
        lines = [ln.split() for ln in fh.read().splitlines() if ln.strip()]
    first = [list(map(float, ln)) for ln in lines[1:1 + nr0]]
    second = [list(map(float, ln)) for ln in lines[2 + nr0:2 + 2 * nr0]]
    chrd = 0.166
    a = np.array(first)
    b = np.array(second)
    out = {
        "rtim": a[:, 0] * 0.7, "delrt": a[:, 1].copy(), "hrt": chrd * a[:, 2] / H_0 / WF,
        "delr": b[:, 0].copy(), "timed": b[:, 1] * 0.7, "hrd": chrd * b[:, 2] / H_0 / WF,
    }
    return {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in out.items()}


def dnst_calc_at(arr, sl):
    """CONC then DNST_CALC on the lattice positions sl (C/Prop.f90:28-59,
    187-198): H2O = 1 - sum of the others (all renormalised when that sum
    exceeds 1), rho = (1 + p)/T * W_mix/W0, at least 1e-20."""
    P, T, DN, GF = arr["p"], arr["t"], arr["dnst"], arr["gf"]
    y = GF[sl]
    s = y[..., IYF - 1] + y[..., IYO2 - 1] + y[..., IYCO2 - 1] + y[..., IYN2 - 1]
    over = s > 1.0
    if np.any(over):
        s2 = s + y[..., IYH2O - 1]
        for q in (IYF, IYO2, IYCO2, IYN2, IYH2O):
            y[..., q - 1] = np.where(over, y[..., q - 1] / s2, y[..., q - 1])
    y[..., IYH2O - 1] = np.where(over, y[..., IYH2O - 1], 1.0 - s)
    GF[sl] = y
    y = GF[sl]
    wt = 1.0 / (y[..., IYF - 1] / WF + y[..., IYCO2 - 1] / WCO2 + y[..., IYH2O - 1] / WH2O +
                y[..., IYO2 - 1] / WO2 + y[..., IYN2 - 1] / WN2)
    DN[sl] = np.maximum((1.0 + P[sl]) / T[sl] * wt / WT0, 1e-20)


def ther_at(arr, c, sl):
    """THER on the lattice positions sl (C/Prop.f90:130-154): eps >= 1e-12,
    mu_t = C_mu Re eps_r/(2 tau) k^2/eps rho, lambda = 1, c_p = CPF1,
    Gamma = 2/Re (mu_t + 1)."""
    GF, DN = arr["gf"], arr["dnst"]
    GF[sl + (IEPS - 1,)] = np.maximum(GF[sl + (IEPS - 1,)], 1.0e-12)
    kk = GF[sl + (IK - 1,)]
    ee = GF[sl + (IEPS - 1,)]
    arr["tmu"][sl] = CMU * c.reyg * c.epsr / (2.0 * c.taur) * kk ** 2 / ee * DN[sl]
    arr["glam"][sl] = 1.0
    arr["gcp"][sl] = CPF1
    arr["gama"][sl] = 2.0 / c.reyg * (arr["tmu"][sl] + 1.0)


class Case:
    """Host state of one case: ``ctrl`` (GfmCtrl) and ``arrays`` (dict of
    Fortran-ordered numpy arrays on the doubled lattice)."""

    def __init__(self, ctrl, arrays, spec):
        self.ctrl = ctrl
        self.arrays = arrays
        self.spec = spec

    @property
    def shape(self):
        return (self.ctrl.mp, self.ctrl.np, self.ctrl.lp)

    @property
    def ncells(self):
        return self.ctrl.ncells

    def clone(self):
        from .abi import copy_arrays
        return Case(self.ctrl.copy(), copy_arrays(self.arrays), self.spec)


def build(spec: CaseSpec) -> Case:
    x, dx = lattice_coord(spec.faces_x)
    r, dr = lattice_coord(spec.faces_y)
    z, dz = lattice_coord(spec.faces_z)
    mp, np_, lp = len(x), len(r), len(z)
    shp = (mp, np_, lp)
    F = lambda v=0.0: np.full(shp, v, dtype=np.float64, order="F")

    # ---- cell codes (GRID3) ----
    ib = np.zeros(shp, dtype=np.int16, order="F")
    cen = ib[1::2, 1::2, 1::2]  # view on centres, index c -> I = 2c+2
    cen[0, :, :] = 1; cen[-1, :, :] = 1
    cen[:, 0, :] = 1; cen[:, -1, :] = 1
    cen[:, :, 0] = 1; cen[:, :, -1] = 1
    if spec.exit_face != "x+":
        raise ValueError(spec.exit_face)
    if spec.exit_yz is None:
        cen[-1, 1:-1, 1:-1] = 3
    else:
        y0, y1, z0, z1 = spec.exit_yz
        yc = 0.5 * (np.asarray(spec.faces_y)[:-1] + np.asarray(spec.faces_y)[1:])
        zc = 0.5 * (np.asarray(spec.faces_z)[:-1] + np.asarray(spec.faces_z)[1:])
        jy = np.nonzero((yc > y0) & (yc < y1))[0] + 1
        kz = np.nonzero((zc > z0) & (zc < z1))[0] + 1
        cen[-1, jy[0]:jy[-1] + 1, kz[0]:kz[-1] + 1] = 3
    if spec.melt_bottom_T is not None:
        cen[1:-1, 1:-1, 0] = 4
    for (c0, c1), (d0, d1), (e0, e1), code in spec.half_blocked:
        if code not in (-1, -2):
            raise ValueError(code)
        sub = cen[1 + c0:2 + c1, 1 + d0:2 + d1, 1 + e0:2 + e1]
        sub[sub == 0] = code
    for inl in spec.inlets:
        i1, j1, k1, i2, j2, k2 = inl.box
        for I in range(i1, i2 + 1, 2):
            for J in range(j1, j2 + 1, 2):
                for K in range(k1, k2 + 1, 2):
                    ib[I - 1, J - 1, K - 1] = 2
    ncells = propagate_ibcell(ib)

    # ---- inlet mass flows in SI (sets GMFR, as the GUI does for the sbc file) ----
    def cell_area(I, J, K, iu):
        if iu == 1:
            return dr[J - 1] * dz[K - 1]
        if iu == 2:
            return dx[I - 1] * dz[K - 1]
        return dx[I - 1] * dr[J - 1]

    taken = set()
    gmfr = 0.0
    for inl in spec.inlets:
        i1, j1, k1, i2, j2, k2 = inl.box
        w = mixture_w([inl.yf, inl.yo2, inl.yco2, inl.yn2, inl.yh2o])
        rho = PG0 * w / (RU * inl.T)
        for I in range(i1, i2 + 1, 2):
            for J in range(j1, j2 + 1, 2):
                for K in range(k1, k2 + 1, 2):
                    if (I, J, K) in taken:
                        continue
                    taken.add((I, J, K))
                    gmfr += rho * abs(inl.uvw[abs(inl.iu) - 1]) * cell_area(I, J, K, abs(inl.iu))

    # ---- SETUP: reference scales (C/Setup.f90:542-731) ----
    dnst0 = PG0 * WT0 / RU / T0
    ug0 = gmfr / dnst0 / 1.0
    tk0 = CTEA * ug0 ** 2
    epsi0 = CTDA * tk0 ** 2
    reyg = 2.0 * dnst0 * ug0 * 1.0 / GMU_REF
    c = GfmCtrl()
    c.mp, c.np, c.lp = mp, np_, lp
    c.nr0 = 100
    c.ncells = ncells
    c.sweep_order = 1
    c.precision = 0
    c.want_pre_residuals = 1
    c.want_irc = 0
    c.i1st_rad_done = 0
    c.isothermal_walls = 0
    c.gray_radiation = 0
    c.itr_gas = 0
    c.itr_gend = spec.maxgi
    c.maxgi = spec.maxgi
    c.bgcon = spec.bgcon
    c.factor = 1.0
    c.fadj = 0.0
    c.t0 = T0
    c.trdg = TL0 / T0
    c.h_0 = H_0
    c.r0 = c.area0 = c.vol0 = 1.0
    c.reyg = reyg
    c.reyg2 = reyg / 2.0
    c.eun = PG0 / (dnst0 * ug0 ** 2)
    c.epsr = tk0 / ug0 ** 2
    c.taur = epsi0 * 1.0 / (tk0 * ug0)
    c.ug0 = ug0
    c.dnst0 = dnst0
    c.qr0 = H_0 * dnst0 * ug0 / 1.0
    c.q0 = spec.q0 / H_0
    c.gmfr = gmfr
    c.tk0 = tk0
    c.epsi0 = epsi0
    g = np.array([0.0, 0.0, -9.80621]) / ug0 / ug0 * 1.0
    for q in range(3):
        c.gra[q] = g[q]
    c.p_ref = 0.0
    c.t_mx = 3.0e3 / T0
    c.t_mn = 2.0e2 / T0
    c.u_mx = 1.0e3 / ug0
    c.u_mn = -c.u_mx
    c.p_mx, c.p_mn = 0.2, -0.6
    c.f_mx, c.f_mn = 1.5, 0.8
    for q in range(21):
        c.rfl[q] = RFL.get(q + 1, RFL_GENERAL)
        c.rf[q] = RF.get(q + 1, RF_GENERAL)
    c.rf_hsource = RF_HSOURCE
    for q in range(20):
        c.ntimesg[q] = NTIMESG[q]
    for q in range(2):
        c.bv_f[q] = spec.bv_f[q]
    for q in range(6):
        c.ba_f[q] = spec.ba_f[q]
    c.sigma_sb = 5.668e-8
    c.h_g = 10.0
    c.kappa_gray = spec.gray_kappa
    c.gray_radiation = 1 if spec.gray_kappa > 0.0 else 0
    c.isothermal_walls = 1 if spec.t_wall is not None else 0
    c.t_wall = (spec.t_wall if spec.t_wall is not None else spec.t_init) / T0

    arr = {"x": x, "dx": dx, "r": r, "dr": dr, "rs": np.ones(np_), "z": z, "dz": dz, "ibcell": ib}
    for k in FIELD3:
        arr[k] = F(0.0)
    arr["dnst"][:] = 1.0
    arr["t"][:] = 1.0
    arr["theta"][:] = 1.0
    arr["ug"] = np.zeros(shp + (3,), dtype=np.float64, order="F")
    arr["gf"] = np.zeros(shp + (9,), dtype=np.float64, order="F")
    arr.update(load_kinetic(c.nr0))
    P, T, DN, TH, TMU, GCP, GLAM, GAMA = (arr[k] for k in ("p", "t", "dnst", "theta", "tmu", "gcp", "glam", "gama"))
    UG, GF = arr["ug"], arr["gf"]

    def dnst_calc(sl):
        dnst_calc_at(arr, sl)

    def ther(sl):
        ther_at(arr, c, sl)

    def tke(sl):
        g = UG[sl + (0,)] ** 2 + UG[sl + (1,)] ** 2 + UG[sl + (2,)] ** 2
        GF[sl + (IK - 1,)] = np.maximum(g, 1.0e-6)
        GF[sl + (IEPS - 1,)] = GF[sl + (IK - 1,)] ** 2

    def t2h(sl):
        y5 = [GF[sl + (q - 1,)] for q in CHEM]
        GF[sl + (IH - 1,)] = t_to_enth(y5, T[sl])

    # ---- SBC: inlet cells (C/Sbc.f90:67-269) ----
    fr_f = fr_ox = fr_n2 = 0.0
    first_inlet = None
    ibw = ib.copy(order="F")
    for inl in spec.inlets:
        i1, j1, k1, i2, j2, k2 = inl.box
        u0 = np.array(inl.uvw) / ug0
        for I in range(i1, i2 + 1, 2):
            for J in range(j1, j2 + 1, 2):
                for K in range(k1, k2 + 1, 2):
                    if ibw[I - 1, J - 1, K - 1] != 2:
                        continue
                    ibw[I - 1, J - 1, K - 1] = 20
                    if first_inlet is None:
                        first_inlet = (I, J, K)
                    s = (I - 1, J - 1, K - 1)
                    sl = tuple(slice(q, q + 1) for q in s)
                    P[s] = c.p_ref
                    T[s] = inl.T / T0
                    GF[s + (IYF - 1,)] = inl.yf
                    GF[s + (IYCO2 - 1,)] = inl.yco2
                    GF[s + (IYH2O - 1,)] = inl.yh2o
                    GF[s + (IYO2 - 1,)] = inl.yo2
                    GF[s + (IYN2 - 1,)] = inl.yn2
                    dnst_calc(sl)
                    UG[s + (0,)], UG[s + (1,)], UG[s + (2,)] = u0
                    tke(sl)
                    ther(sl)
                    t2h(sl)
                    g0 = TH[s] * DN[s] * dnst0 * abs(u0[abs(inl.iu) - 1]) * ug0 * cell_area(I, J, K, abs(inl.iu))
                    fr_f += g0 * GF[s + (IYF - 1,)]
                    fr_ox += g0 * GF[s + (IYO2 - 1,)]
                    fr_n2 += g0 * GF[s + (IYN2 - 1,)]
        _sbc_fill_block(inl, ibw, x, r, z, dx, dr, dz, P, T, GF, UG, DN, TH, TMU, GCP, GLAM, GAMA, dnst_calc)
    gfin = fr_f + fr_ox + fr_n2
    c.gfin0 = gfin / gmfr
    c.i_ref, c.j_ref, c.k_ref = first_inlet
    yo2_out = (fr_ox - 2 * fr_f * WO2 / WF) / gfin
    yco2_out = (fr_f * WCO2 / WF) / gfin
    yh2o_out = (2 * fr_f * WH2O / WF) / gfin
    yn2_out = fr_n2 / gfin

    # ---- gen_initial_guess (C/Initsv.f90:97-162) ----
    cen_ib = ib[1::2, 1::2, 1::2]
    sl_c = (slice(1, None, 2), slice(1, None, 2), slice(1, None, 2))
    notin = cen_ib != 2
    for q, v in ((IYF, 0.0), (IYO2, yo2_out), (IYCO2, yco2_out), (IYH2O, yh2o_out), (IYN2, yn2_out)):
        view = GF[sl_c + (q - 1,)]
        view[notin] = v
        GF[sl_c + (q - 1,)] = view
    wall = cen_ib == 1
    Tc = T[sl_c]
    Tc[wall] = (spec.t_wall if spec.t_wall is not None else spec.t_init) / T0
    T[sl_c] = Tc
    y5 = [GF[sl_c + (q - 1,)] for q in CHEM]
    hc = GF[sl_c + (IH - 1,)]
    hc[wall] = t_to_enth(y5, T[sl_c])[wall]
    GF[sl_c + (IH - 1,)] = hc
    opn = notin & ~wall
    for q in range(3):
        v = UG[sl_c + (q,)]; v[opn] = 0.0; UG[sl_c + (q,)] = v
    v = TH[sl_c]; v[opn] = 1.0; TH[sl_c] = v
    v = T[sl_c]; v[opn & (cen_ib != 4)] = spec.t_init / T0; T[sl_c] = v
    rng = np.random.default_rng(spec.seed)
    if spec.melt_bottom_T is not None:
        xc = x[1::2]
        tb = np.array([spec.melt_bottom_T(xx) for xx in xc]) / T0
        v = T[sl_c]
        m4 = cen_ib == 4
        v[m4] = np.broadcast_to(tb[:, None, None], v.shape)[m4]
        T[sl_c] = v
    # TKE from the (zero) velocity, then the seeded initial turbulence
    g = UG[sl_c + (0,)] ** 2 + UG[sl_c + (1,)] ** 2 + UG[sl_c + (2,)] ** 2
    kk = GF[sl_c + (IK - 1,)]; ee = GF[sl_c + (IEPS - 1,)]
    kk[opn] = np.maximum(g, 1.0e-6)[opn]
    ee[opn] = kk[opn] ** 2
    if spec.k_init is not None:
        kk[opn] = spec.k_init / tk0
    if spec.eps_init is not None:
        ee[opn] = spec.eps_init / epsi0
    if spec.noise > 0.0:
        n_open = int(np.count_nonzero(opn))
        Tv = T[sl_c]
        tnz = 1.0 + spec.noise * rng.uniform(-1.0, 1.0, n_open)
        knz = 1.0 + spec.noise * rng.uniform(-1.0, 1.0, n_open)
        enz = 1.0 + spec.noise * rng.uniform(-1.0, 1.0, n_open)
        m = opn & (cen_ib != 4)
        Tv[opn] = Tv[opn] * np.where(m[opn], tnz, 1.0)
        T[sl_c] = Tv
        kk[opn] = kk[opn] * knz
        ee[opn] = ee[opn] * enz
    GF[sl_c + (IK - 1,)] = kk
    GF[sl_c + (IEPS - 1,)] = ee
    y5 = [GF[sl_c + (q - 1,)] for q in CHEM]
    hc = GF[sl_c + (IH - 1,)]
    h_new = t_to_enth(y5, T[sl_c])
    hc[opn] = h_new[opn]
    hc[cen_ib == 4] = h_new[cen_ib == 4]
    GF[sl_c + (IH - 1,)] = hc
    # dnst_calc and ther on the open cells
    dn_all = DN[sl_c].copy(); tm_all = TMU[sl_c].copy(); ga_all = GAMA[sl_c].copy(); gc_all = GCP[sl_c].copy(); gl_all = GLAM[sl_c].copy()
    gf_all = GF[sl_c].copy()
    dnst_calc(sl_c)
    ther(sl_c)
    for arrv, old in ((DN, dn_all), (TMU, tm_all), (GAMA, ga_all), (GCP, gc_all), (GLAM, gl_all)):
        v = arrv[sl_c]
        v[~opn] = old[~opn]
        arrv[sl_c] = v
    v = GF[sl_c]
    v[~opn] = gf_all[~opn]
    GF[sl_c] = v
    case = Case(c, arr, spec)
    case.meta = dict(fr_f=fr_f, fr_ox=fr_ox, fr_n2=fr_n2, gmfr=gmfr, ug0=ug0, ncells=ncells,
                     air_fuel=(fr_ox + fr_n2) / fr_f if fr_f > 0 else float("inf"))
    return case


def _sbc_fill_block(inl, ibw, x, r, z, dx, dr, dz, P, T, GF, UG, DN, TH, TMU, GCP, GLAM, GAMA, dnst_calc):
    """SBC1/SBC3 fills of the face, edge and neighbour positions of an inlet
    block (C/Sbc.f90:179-267, :463-522)."""
    i1, j1, k1, i2, j2, k2 = inl.box
    iu = inl.iu
    ibc = lambda I, J, K: int(ibw[I - 1, J - 1, K - 1])

    def sbc1(a, b, c_, f1):
        f2 = 1.0 - f1
        A = tuple(q - 1 for q in a); B = tuple(q - 1 for q in b); Cc = tuple(q - 1 for q in c_)
        P[Cc] = f1 * P[A] + f2 * P[B]
        T[Cc] = f1 * T[A] + f2 * T[B]
        for L in range(9):
            if L < 3:
                UG[Cc + (L,)] = f1 * UG[A + (L,)] + f2 * UG[B + (L,)]
            GF[Cc + (L,)] = f1 * GF[A + (L,)] + f2 * GF[B + (L,)]
        dnst_calc(tuple(slice(q, q + 1) for q in Cc))
        for arr in (TH, TMU, GCP, GLAM, GAMA):
            arr[Cc] = f1 * arr[A] + f2 * arr[B]

    def sbc3(src, dst):
        S = tuple(q - 1 for q in src); Dd = tuple(q - 1 for q in dst)
        for L in range(3):
            UG[Dd + (L,)] = UG[S + (L,)]
        P[Dd] = P[S]; T[Dd] = T[S]
        for L in range(9):
            GF[Dd + (L,)] = GF[S + (L,)]
        for arr in (DN, TH, TMU, GLAM, GCP, GAMA):
            arr[Dd] = arr[S]

    mp, np_, lp = ibw.shape
    if abs(iu) == 1 or iu == 1:
        I = i1
        if k2 > k1:
            for K in range(k1 + 1, k2, 2):
                for J in range(j1, j2 + 1, 2):
                    sbc1((I, J, K + 1), (I, J, K - 1), (I, J, K), (z[K - 1] - z[K - 2]) / dz[K - 1])
        if j2 > j1:
            for J in range(j1 + 1, j2, 2):
                for K in range(k1, k2 + 1):
                    sbc1((I, J + 1, K), (I, J - 1, K), (I, J, K), (r[J - 1] - r[J - 2]) / dr[J - 1])
        for J in range(j1, j2 + 1):
            for K in range(k1, k2 + 1):
                if I < mp and ibc(I + 1, J, K) == 2:
                    sbc3((I, J, K), (I + 1, J, K))
                if I > 2 and ibc(I - 1, J, K) == 2:
                    sbc3((I, J, K), (I - 1, J, K))
    elif abs(iu) == 2:
        J = j1
        if k2 > k1:
            for K in range(k1 + 1, k2, 2):
                for I in range(i1, i2 + 1, 2):
                    sbc1((I, J, K + 1), (I, J, K - 1), (I, J, K), (z[K - 1] - z[K - 2]) / dz[K - 1])
        if i2 > i1:
            for I in range(i1 + 1, i2, 2):
                for K in range(k1, k2 + 1):
                    sbc1((I + 1, J, K), (I - 1, J, K), (I, J, K), (x[I - 1] - x[I - 2]) / dx[I - 1])
        for I in range(i1, i2 + 1):
            for K in range(k1, k2 + 1):
                if J < np_ and ibc(I, J + 1, K) == 2:
                    sbc3((I, J, K), (I, J + 1, K))
                if J > 2 and ibc(I, J - 1, K) == 2:
                    sbc3((I, J, K), (I, J - 1, K))
    else:
        K = k1
        if j2 > j1:
            for J in range(j1 + 1, j2, 2):
                for I in range(i1, i2 + 1, 2):
                    sbc1((I, J + 1, K), (I, J - 1, K), (I, J, K), (r[J - 1] - r[J - 2]) / dr[J - 1])
        if i2 > i1:
            for I in range(i1 + 1, i2, 2):
                for J in range(j1, j2 + 1):
                    sbc1((I + 1, J, K), (I - 1, J, K), (I, J, K), (x[I - 1] - x[I - 2]) / dx[I - 1])
        for I in range(i1, i2 + 1):
            for J in range(j1, j2 + 1):
                if K < lp and ibc(I, J, K + 1) == 2:
                    sbc3((I, J, K), (I, J, K + 1))
                if K > 2 and ibc(I, J, K - 1) == 2:
                    sbc3((I, J, K), (I, J, K - 1))


# ----------------------------------------------------------------------
# The benchmark furnaces
# ----------------------------------------------------------------------
def cell_I(c):
    """Lattice index of 0-based interior cell c (first interior centre is 4)."""
    return 4 + 2 * c


def uniform_faces(length, n):
    return np.linspace(0.0, length, n + 1)


def stretched_faces(length, n, ratio):
    """Faces clustered toward both ends with geometric growth ``ratio`` per
    cell toward the middle (n even)."""
    half = n // 2
    w = ratio ** np.arange(half)
    w = w / w.sum() * (length / 2.0)
    left = np.concatenate([[0.0], np.cumsum(w)])
    right = length - left[::-1]
    return np.concatenate([left, right[1:]])


def air_velocity(u_air0, fuel_u, fuel_u0, air_t, air_t0, excess_air, excess_air0=0.10):
    """Air-port velocity of an operating point: the nominal velocity scaled
    with the fuel velocity, with (1 + excess air) and with the preheat
    temperature (constant mass ratio at ideal-gas density).  The nominal
    port sizes of the furnaces cannot pass a stoichiometric air flow at the
    stated velocities, so excess air is applied relative to the nominal
    point rather than to the stoichiometric ratio."""
    return u_air0 * (fuel_u / fuel_u0) * ((1.0 + excess_air) / (1.0 + excess_air0)) * (air_t / air_t0)


def furnace_small(nx=48, ny=24, nz=16, seed=1, noise=0.01, maxgi=300, bgcon=1.0e-12, walls="isothermal",
                  gray_kappa=0.3, fuel_u=30.0, air_t=1100.0, excess_air=0.10, air_port=(0.3, 0.2), air_u=None,
                  reacting=True, half_blocked=(), bv_f=(1.0, 1.0), ba_f=(1.0,) * 6):
    """3.0 x 1.2 x 0.8 m chamber, uniform cells, one CH4 port (0.1x0.1 m,
    30 m/s, 300 K) centred on the x=0 wall inside an air port (0.3x0.2 m,
    10 m/s, 1100 K), full-face exit at x=3 m; uniform initial field
    u=0, T=1500 K, k=eps=1e-2 with +-1 % seeded noise; walls isothermal at
    1500 K and a gray radiation source with kappa=0.3 1/m (walls="adiabatic",
    gray_kappa=0 give the reference's behaviour before its first radiation
    pass).  For grids other than 48x24x16 the ports keep their physical
    size.  fuel_u, air_t, excess_air set an operating point (air velocity
    from air_velocity, nominal 10 m/s, or air_u).  air_port (y, z extent in
    m) sizes the air port; (1.2, 0.8) makes it the whole x=0 wall around
    the fuel port (a coflow chamber).  reacting=False sends air at 300 K
    through the fuel port (no reaction).  half_blocked, bv_f, ba_f: cells
    of code -1/-2 and their &GRD factors (CaseSpec)."""
    Lx, Ly, Lz = 3.0, 1.2, 0.8
    u_air = air_velocity(10.0, fuel_u, 30.0, air_t, 1100.0, excess_air)
    dy, dz = Ly / ny, Lz / nz

    def span(centre, width, d, n):
        a = int(math.floor((centre - width / 2) / d + 0.5))
        b = int(math.floor((centre + width / 2) / d + 0.5)) - 1
        b = max(b, a)
        return max(a, 0), min(b, n - 1)

    fy = span(0.6, 0.1, dy, ny); fz = span(0.4, 0.1, dz, nz)
    ay = span(0.6, air_port[0], dy, ny); az = span(0.4, air_port[1], dz, nz)
    if air_u is not None:
        u_air = air_u
    fy_f, fy_o2, fy_n2 = (1.0, 0.0, 0.0) if reacting else (0.0, 0.233, 0.767)
    fuel = Inlet(box=(2, cell_I(fy[0]), cell_I(fz[0]), 2, cell_I(fy[1]), cell_I(fz[1])), iu=1,
                 T=300.0, yf=fy_f, yo2=fy_o2, yn2=fy_n2, uvw=(fuel_u, 0.0, 0.0))
    air = Inlet(box=(2, cell_I(ay[0]), cell_I(az[0]), 2, cell_I(ay[1]), cell_I(az[1])), iu=1,
                T=air_t, yf=0.0, yo2=0.233, yn2=0.767, uvw=(u_air, 0.0, 0.0))
    spec = CaseSpec(name=f"furnace {nx}x{ny}x{nz}", faces_x=uniform_faces(Lx, nx), faces_y=uniform_faces(Ly, ny),
                    faces_z=uniform_faces(Lz, nz), inlets=[fuel, air], t_init=1500.0, k_init=1.0e-2,
                    eps_init=1.0e-2, noise=noise, seed=seed, maxgi=maxgi, bgcon=bgcon,
                    t_wall=1500.0 if walls == "isothermal" else None, gray_kappa=gray_kappa,
                    half_blocked=list(half_blocked), bv_f=tuple(bv_f), ba_f=tuple(ba_f))
    return build(spec)


def furnace_large(nx=192, ny=96, nz=64, seed=2, noise=0.01, maxgi=2000, bgcon=1.0e-5, walls="isothermal",
                  gray_kappa=0.3, exit_port=None, fuel_u=25.0, air_t=1200.0, excess_air=0.10):
    """7.0 x 3.0 x 1.6 m furnace, y-faces stretched toward the burner walls
    (ratio 1.05), two opposed CH4 burner pairs on the y walls at x=1.5 m and
    x=4.0 m (fuel 25 m/s 300 K, air 12 m/s 1200 K; the air port is sized
    for ~10 % excess air), glass bottom (IBCELL 4) at 1450 K +- 20 K
    sinusoidal in x, full-face exit at x=7 m.  fuel_u, air_t, excess_air set
    an operating point (air velocity from air_velocity, nominal 12 m/s)."""
    Lx, Ly, Lz = 7.0, 3.0, 1.6
    fx = uniform_faces(Lx, nx)
    fy = stretched_faces(Ly, ny, 1.05)
    fzf = uniform_faces(Lz, nz)
    dxc, dzc = Lx / nx, Lz / nz
    rng = np.random.default_rng(seed)
    phase = rng.uniform(0.0, 2.0 * math.pi)
    u_air = air_velocity(12.0, fuel_u, 25.0, air_t, 1200.0, excess_air)
    inlets = []
    # fuel 1x1 cell at the port centre, air port 9 (x) by 10 (z) cells
    for xb in (1.5, 4.0):
        ci = int(round(xb / dxc)) - 1
        ck = int(round(0.5 * Lz / dzc)) - 1
        for side, jw, v in ((+1, 2, 1.0), (-1, 2 * ny + 4, -1.0)):
            fuel = Inlet(box=(cell_I(ci), jw, cell_I(ck), cell_I(ci), jw, cell_I(ck)), iu=2 * side,
                         T=300.0, yf=1.0, yo2=0.0, yn2=0.0, uvw=(0.0, fuel_u * v, 0.0))
            air = Inlet(box=(cell_I(ci - 4), jw, cell_I(ck - 4), cell_I(ci + 4), jw, cell_I(ck + 5)), iu=2 * side,
                        T=air_t, yf=0.0, yo2=0.233, yn2=0.767, uvw=(0.0, u_air * v, 0.0))
            inlets += [fuel, air]
    spec = CaseSpec(name=f"furnace {nx}x{ny}x{nz} stretched", faces_x=fx, faces_y=fy, faces_z=fzf, inlets=inlets,
                    t_init=1500.0, k_init=1.0e-2, eps_init=1.0e-2, noise=noise, seed=seed, maxgi=maxgi, bgcon=bgcon,
                    melt_bottom_T=lambda xx: 1450.0 + 20.0 * math.sin(2.0 * math.pi * xx / Lx + phase),
                    t_wall=1600.0 if walls == "isothermal" else None, gray_kappa=gray_kappa, exit_yz=exit_port)
    return build(spec)


def from_grid(grid, gmfr):
    """A case on a reference grid file (gridfile.read_gd): the lattice,
    the cell codes and the &GRD factors of the file, SETUP's scales for the
    mass flow gmfr (kg/s, the sbc file's GMFR), uniform fields (T = T0,
    rho = rho0).  The inlet fills of SBC are not applied: this holds the
    geometry on which a saved state (restart.apply_records, the rs record)
    is read, e.g. for a radiation pass on a reference run's state."""
    from .abi import GfmCtrl
    from .gridfile import grd_factors
    mp, np_, lp = grid.shape
    shp = (mp, np_, lp)
    dnst0 = PG0 * WT0 / RU / T0
    ug0 = gmfr / dnst0 / 1.0
    tk0 = CTEA * ug0 ** 2
    epsi0 = CTDA * tk0 ** 2
    c = GfmCtrl()
    c.mp, c.np, c.lp, c.nr0 = mp, np_, lp, 100
    c.ncells = grid.ncells
    c.sweep_order, c.precision = 1, 0
    c.t0, c.trdg, c.h_0 = T0, TL0 / T0, H_0
    c.r0 = c.area0 = c.vol0 = 1.0
    c.reyg = 2.0 * dnst0 * ug0 / GMU_REF
    c.reyg2 = c.reyg / 2.0
    c.eun = PG0 / (dnst0 * ug0 ** 2)
    c.epsr = tk0 / ug0 ** 2
    c.taur = epsi0 / (tk0 * ug0)
    c.ug0, c.dnst0, c.gmfr, c.tk0, c.epsi0 = ug0, dnst0, gmfr, tk0, epsi0
    c.qr0 = H_0 * dnst0 * ug0
    bv, ba = grd_factors(grid.grd)
    for q in range(2):
        c.bv_f[q] = bv[q]
    for q in range(6):
        c.ba_f[q] = ba[q]
    c.sigma_sb = 5.668e-8
    c.factor = 1.0
    arr = {"x": grid.x.copy(), "dx": grid.dx.copy(), "r": grid.r.copy(), "dr": grid.dr.copy(), "rs": np.ones(np_),
           "z": grid.z.copy(), "dz": grid.dz.copy(), "ibcell": np.array(grid.ibcell, dtype=np.int16, order="F")}
    for k in FIELD3:
        arr[k] = np.zeros(shp, order="F")
    arr["dnst"][:] = 1.0
    arr["t"][:] = 1.0
    arr["theta"][:] = 1.0
    arr["ug"] = np.zeros(shp + (3,), order="F")
    arr["gf"] = np.zeros(shp + (9,), order="F")
    arr.update(load_kinetic(c.nr0))
    return Case(c, arr, None)
