"""Checks that pin the oracle's stage functions which the closed forms of
tests/test_oracle.py leave open: each stage runs alone on a prepared state
(oracle.probe) and is compared with an invariant or a closed form that
does not restate its formula, chosen so that a plausible slip (a dropped
term, a wrong sign, index or factor) fails.

  p' equation, velocity correction   after the correction every cell's mass
  (C/Gslv3v.f90:43-101,149-204,      imbalance / Eu equals its p'-equation
   C/Source.f90:41-101)              residual, at any sweep count: the p'
                                     coefficients, SOURCE(5)'s b and the
                                     correction formula are one algebra.
  INTP, INTPV (C/Intp.f90:32-272)    a field linear in (x, y, z) comes out
                                     exact at every position they write, on
                                     a stretched grid; at an inlet position
                                     between an inlet and an open cell the
                                     inlet side's value is taken.
  GSLVX1 + SRC1 (C/Gslv3x.f90:26-    uniform flow: a_p = sum a_nb, the
   111, C/Source.f90:664-809)        pseudo-velocity equals the flow, the
                                     source is buoyancy times the staggered
                                     volume; linear fields: the normal and
                                     cross viscous terms equal d/dx_j(Gamma
                                     du_j/dx_i) of the fields.
  q_conv (C/Qwall.f90:1419-1467)     h_g below the k threshold; the cap of
                                     50; 1/h_t linear in ln(ds) with slope
                                     1/(c_p kappa C_mu^1/4 rho sqrt(k)).
  species sources (C/Source.f90:     fuel sink = -crate, CO2 = crate
   104-266,621-654)                  W_CO2/W_F, O2 = -2 W_O2/W_F crate,
                                     N2 = 0, and the element balance of
                                     CH4 + 2 O2 -> CO2 + 2 H2O closes.
  enthalpy-source history           with a known history field the relaxed
  (C/Source.f90:511-520)             source is rf_h raw + (1 - rf_h) history,
                                     stored as the next history.
  CONC, DNST_CALC, THER              the ideal-gas law in SI units, mass
  (C/Prop.f90:28-59,130-198)         fractions summing to one, mu_t =
                                     C_mu rho k^2/eps in SI over the
                                     reference viscosity.
"""
import math

import numpy as np
import pytest

import oracle
from gfm_amd import case as cs

IYF, IYO2, IH, IK, IEPS, IYCO2, IYN2, IYH2O = cs.IYF, cs.IYO2, cs.IH, cs.IK, cs.IEPS, cs.IYCO2, cs.IYN2, cs.IYH2O


def _stretched():
    return cs.furnace_large(nx=32, ny=24, nz=12, seed=3)


def _uniform():
    return cs.furnace_small(nx=16, ny=12, nz=10, seed=3)


def _cell_slices(c):
    mp, np_, lp = c.shape
    return (slice(3, mp - 2, 2), slice(3, np_ - 2, 2), slice(3, lp - 2, 2))


# ---------------------------------------------------------------------
# p' equation and velocity correction
# ---------------------------------------------------------------------
def _imbalance_over_eu(c):
    """(mass in - mass out)/Eu of every cell of the p' box from the current
    face velocities, with the face theta*rho and the cell's face areas."""
    A = c.arrays
    th, dn, ug, dx, dr, dz, rs = A["theta"], A["dnst"], A["ug"], A["dx"], A["dr"], A["dz"], A["rs"]
    mp, np_, lp = c.shape
    I, J, K = np.arange(3, mp - 2, 2), np.arange(3, np_ - 2, 2), np.arange(3, lp - 2, 2)
    ax = (dr[J] * rs[J])[None, :, None] * dz[K][None, None, :]
    ay = dx[I][:, None, None] * dz[K][None, None, :]
    az = dx[I][:, None, None] * (dr[J] * rs[J])[None, :, None]
    g = th * dn

    def F(di, dj, dk, q):
        return (g * ug[..., q])[3 + di:mp - 2 + di:2, 3 + dj:np_ - 2 + dj:2, 3 + dk:lp - 2 + dk:2]

    r = (F(-1, 0, 0, 0) - F(1, 0, 0, 0)) * ax + (F(0, -1, 0, 1) - F(0, 1, 0, 1)) * ay + (F(0, 0, -1, 2) - F(0, 0, 1, 2)) * az
    return r / c.ctrl.eun


def _pcorr_residual(c, s):
    ap, as_, bs, x = s["ap"], s["as_"], s["bs"], s["fz"]
    mp, np_, lp = c.shape
    sl = _cell_slices(c)

    def sh(di, dj, dk):
        return x[3 + di:mp - 2 + di:2, 3 + dj:np_ - 2 + dj:2, 3 + dk:lp - 2 + dk:2]

    a = as_[sl]
    return (a[..., 0, 0] * sh(-2, 0, 0) + a[..., 0, 1] * sh(2, 0, 0) + a[..., 1, 0] * sh(0, -2, 0) + a[..., 1, 1] * sh(0, 2, 0)
            + a[..., 2, 0] * sh(0, 0, -2) + a[..., 2, 1] * sh(0, 0, 2) + bs[sl] - ap[sl] * x[sl])


@pytest.mark.parametrize("grid", ["uniform", "stretched"])
@pytest.mark.parametrize("mode", ["reference", "tolerance"])
def test_velocity_correction_matches_pcorr_residual(grid, mode):
    c = _uniform() if grid == "uniform" else _stretched()
    oracle.initsv_tail(c)
    oracle.sphase(c, 4)
    if mode == "tolerance":
        c.ctrl.pcorr_mode = 1
        c.ctrl.pcorr_tol = 0.0
        c.ctrl.pcorr_max_sweeps = 40
        c.ctrl.pcorr_block_corr = 1
    c.ctrl.itr_gas += 1
    oracle.probe(c, "vel_correct")
    s = oracle.scratch(c)
    rp = _pcorr_residual(c, s)
    rc = _imbalance_over_eu(c)
    opn = c.arrays["ibcell"][_cell_slices(c)] <= 0
    b = s["bs"][_cell_slices(c)]
    scale = np.abs(b[opn]).max()
    err = np.abs(rc - rp)[opn].max() / scale
    print(grid, mode, "max |imbalance/Eu - p' residual| / max|b|", err, flush=True)
    assert scale > 0.0
    assert err < 1e-10
    # b is SOURCE(5)'s imbalance of the starred velocities; a correction
    # scaled by rf (the relaxed a_p in place of the unrelaxed one) would
    # leave the imbalance far from the p' residual
    slip = b + (rc - b) * c.ctrl.rf[0]
    assert np.abs(slip - rp)[opn].max() / scale > 1e-3


# ---------------------------------------------------------------------
# INTP / INTPV on linear fields
# ---------------------------------------------------------------------
def _coords(c):
    A = c.arrays
    mp, np_, lp = c.shape
    X = np.broadcast_to(A["x"][:, None, None], (mp, np_, lp))
    Y = np.broadcast_to(A["r"][None, :, None], (mp, np_, lp))
    Z = np.broadcast_to(A["z"][None, None, :], (mp, np_, lp))
    return X, Y, Z


LINEAR = {  # field: (a0, ax, ay, az) of a0 + ax x + ay y + az z (x, y, z in m)
    "p": (0.01, 0.002, -0.001, 0.003), "t": (3.0, 0.05, 0.02, -0.04), "tmu": (50.0, 2.0, -3.0, 4.0),
    "gcp": (2.0, 0.01, 0.02, 0.03), "glam": (1.0, 0.01, -0.01, 0.02), "gama": (0.02, 1e-3, 2e-3, -1e-3),
    "theta": (0.9, 0.01, -0.01, 0.005),
}
LINEAR_GF = {IYF: (0.05, 0.002, 0.001, -0.003), IYO2: (0.15, -0.004, 0.003, 0.002), IH: (5.0, 0.3, -0.2, 0.1),
             IK: (2.0, 0.1, 0.1, -0.2), IEPS: (3.0, -0.1, 0.2, 0.1), IYCO2: (0.08, 0.003, -0.002, 0.001),
             IYN2: (0.6, 0.001, 0.002, 0.003)}


def _lin(coef, X, Y, Z):
    return coef[0] + coef[1] * X + coef[2] * Y + coef[3] * Z


def _set_linear(c):
    X, Y, Z = _coords(c)
    A = c.arrays
    for f, co in LINEAR.items():
        A[f][...] = _lin(co, X, Y, Z)
    for l, co in LINEAR_GF.items():
        A["gf"][..., l - 1] = _lin(co, X, Y, Z)
    A["gf"][..., IYH2O - 1] = 1.0 - sum(A["gf"][..., q - 1] for q in (IYF, IYO2, IYCO2, IYN2))
    return X, Y, Z


def _intp_positions(c):
    """Positions INTP writes (C/Intp.f90:32-72): every non-centre position
    except walls (IBCELL 1) and exits (3), within the loop ranges."""
    ib = c.arrays["ibcell"]
    mp, np_, lp = c.shape
    w = np.zeros(ib.shape, dtype=bool)
    w[1::2, 1::2, 2:lp - 1:2] = True           # i even, j even, k odd (3..LP-1)
    w[1::2, 2:np_ - 1:2, 1:] = True            # i even, j odd (3..NP-1), k 2..LP
    w[2:mp - 1:2, 1:, 1:] = True               # i odd (3..MP-1), j, k 2..
    return w & (ib != 1) & (ib != 3)


@pytest.mark.parametrize("grid", ["uniform", "stretched"])
def test_intp_exact_on_linear_fields_without_inlets(grid):
    c = _uniform() if grid == "uniform" else _stretched()
    ib = c.arrays["ibcell"]
    ib[ib == 2] = 1  # no inlets: INTP's weight is the geometric one everywhere
    X, Y, Z = _set_linear(c)
    w = _intp_positions(c)
    c.arrays["p"][w] = 123.0  # anything INTP must overwrite
    c.arrays["t"][w] = -7.0
    oracle.probe(c, "intp")
    A = c.arrays
    for f, co in LINEAR.items():
        want = _lin(co, X, Y, Z)[w]
        np.testing.assert_allclose(A[f][w], want, rtol=1e-13, atol=1e-13 * np.abs(want).max(), err_msg=f)
    for l, co in LINEAR_GF.items():
        want = _lin(co, X, Y, Z)[w]
        np.testing.assert_allclose(A["gf"][..., l - 1][w], want, rtol=1e-13, atol=1e-13 * np.abs(want).max(), err_msg=str(l))
    # CONC: H2O closes the mass fractions; DNST_CALC: the ideal-gas law
    ysum = sum(A["gf"][..., q - 1][w] for q in (IYF, IYO2, IYCO2, IYN2, IYH2O))
    np.testing.assert_allclose(ysum, 1.0, rtol=0, atol=1e-14)
    np.testing.assert_allclose(A["dnst"][w], _rho_ideal_gas(c, w) / c.ctrl.dnst0, rtol=1e-12)


def test_intp_takes_inlet_side_value_at_inlet_faces():
    """An inlet position between an inlet cell and an open cell takes the
    inlet side's value (C/Intp.f90:149-157), not the geometric blend."""
    c = _uniform()
    X, Y, Z = _set_linear(c)
    ib = c.arrays["ibcell"]
    oracle.probe(c, "intp")
    mp, np_, lp = c.shape
    n = 0
    for I in range(3, mp - 1, 2):  # x-face loop: neighbours I-1 (i2) and I+1 (i1)
        for J in range(2, np_ + 1):
            for K in range(2, lp + 1):
                if ib[I - 1, J - 1, K - 1] != 2:
                    continue
                a, b = ib[I, J - 1, K - 1], ib[I - 2, J - 1, K - 1]
                if a == 2 and b == 0:
                    src = (I, J - 1, K - 1)
                elif a == 0 and b == 2:
                    src = (I - 2, J - 1, K - 1)
                else:
                    continue
                assert c.arrays["t"][I - 1, J - 1, K - 1] == c.arrays["t"][src]
                n += 1
    assert n > 0


@pytest.mark.parametrize("grid", ["uniform", "stretched"])
def test_intpv_exact_on_linear_velocity(grid):
    c = _uniform() if grid == "uniform" else _stretched()
    X, Y, Z = _coords(c)
    ib = c.arrays["ibcell"]
    co = [(0.2, 0.03, -0.02, 0.01), (-0.1, 0.01, 0.02, -0.03), (0.05, -0.02, 0.01, 0.02)]
    for iu in range(3):
        c.arrays["ug"][..., iu] = _lin(co[iu], X, Y, Z)
    oracle.probe(c, "intpv")
    checked = 0
    for iu in range(3):
        got = c.arrays["ug"][..., iu]
        want = _lin(co[iu], X, Y, Z)
        # positions INTPV leaves alone: walls, inlets, exits; a position next
        # to an inlet along y or z takes the inlet side (not checked here)
        keep = (ib <= 0) | (ib == 4)
        near_inlet = np.zeros(ib.shape, dtype=bool)
        for ax in (1, 2):
            up = np.roll(ib, -1, axis=ax) == 2
            dn = np.roll(ib, 1, axis=ax) == 2
            near_inlet |= up | dn
        m = keep & ~near_inlet
        m[0, :, :] = m[:, 0, :] = m[:, :, 0] = False
        err = np.abs(got - want)[m].max() / np.abs(want[m]).max()
        assert err < 1e-13, (iu, err)
        checked += int(m.sum())
    assert checked > 0


# ---------------------------------------------------------------------
# momentum assembly GSLVX1 + SRC1
# ---------------------------------------------------------------------
def _momentum_case(u_axis=None, grav=False):
    c = _stretched()
    A = c.arrays
    A["theta"][...] = 1.0
    A["dnst"][...] = 1.0
    A["gama"][...] = 0.02
    A["ug"][...] = 0.0
    if u_axis is not None:
        A["ug"][..., u_axis] = 0.4
    if not grav:
        for q in range(3):
            c.ctrl.gra[q] = 0.0
    return c


def _face_box(c, iu):
    """Open faces of component iu (1-based) in GSLVX1's box whose cell
    neighbours two lattice steps either side along every axis are open
    (away from walls, inlets and exits)."""
    ib = c.arrays["ibcell"]
    mp, np_, lp = c.shape
    lo = [4, 4, 4]
    hi = [mp - 2, np_ - 2, lp - 2]
    lo[iu - 1] += 1
    hi[iu - 1] -= 1
    m = np.zeros(ib.shape, dtype=bool)
    m[lo[0] - 1:hi[0]:2, lo[1] - 1:hi[1]:2, lo[2] - 1:hi[2]:2] = True
    m &= ib <= 0
    clear = m.copy()
    for ax in range(3):
        for d in (-3, -2, -1, 1, 2, 3):
            clear &= np.roll(ib, d, axis=ax) <= 0
    return m, clear


@pytest.mark.parametrize("iu", [1, 2, 3])
def test_momentum_uniform_flow(iu):
    c = _momentum_case(u_axis=iu - 1)
    oracle.probe(c, "momentum")
    s = oracle.scratch(c)
    m, clear = _face_box(c, iu)
    ap, as_, bs = s["ap"], s["as_"], s["bs"]
    np.testing.assert_allclose(ap[m], as_[m].sum(axis=(-1, -2)), rtol=1e-13)
    np.testing.assert_allclose(c.arrays["ug"][..., iu - 1][m], 0.4, rtol=1e-13)
    assert np.abs(bs[m]).max() <= 1e-13 * np.abs(ap[m]).max()


def test_momentum_buoyancy_is_rho_g_times_staggered_volume():
    c = _momentum_case(grav=True)
    g = c.ctrl.gra[2]
    assert g < 0.0
    oracle.probe(c, "momentum")
    s = oracle.scratch(c)
    A = c.arrays
    m, clear = _face_box(c, 3)
    X, Y, Z = _coords(c)
    I, J, K = np.nonzero(clear)
    vol = (A["z"][K + 1] - A["z"][K - 1]) * A["dx"][I] * A["dr"][J] * A["rs"][J]
    np.testing.assert_allclose(s["bs"][I, J, K], g * vol, rtol=1e-12)
    # the other components carry no body force
    for iu in (1, 2):
        mm, cl = _face_box(c, iu)
        assert np.abs(s["bs"][cl]).max() <= 1e-14


@pytest.mark.parametrize("case", ["normal_x", "cross_yx", "cross_xy", "cross_zx"])
def test_momentum_viscous_terms_of_linear_fields(case):
    """SRC1 forms d/dx_j (Gamma du_j/dx_i) for face component i; with a
    velocity linear in one coordinate and Gamma linear in another the term
    is the product of the two slopes, times the staggered volume after
    GSLVX1 (C/Gslv3x.f90:107)."""
    c = _momentum_case()
    A = c.arrays
    X, Y, Z = _coords(c)
    slope_u, slope_g = 0.3, 0.05
    # (component solved, component set, its coordinate, Gamma's coordinate)
    spec = {"normal_x": (1, 1, X, X), "cross_yx": (1, 2, X, Y), "cross_xy": (2, 1, Y, X), "cross_zx": (1, 3, X, Z)}
    iu, ju, ucoord, gcoord = spec[case]
    A["ug"][..., ju - 1] = slope_u * ucoord
    A["gama"][...] = 0.02 + slope_g * gcoord
    oracle.probe(c, "momentum")
    s = oracle.scratch(c)
    m, clear = _face_box(c, iu)
    I, J, K = np.nonzero(clear)
    d = [A["dx"][I] if iu != 1 else A["x"][I + 1] - A["x"][I - 1],
         A["dr"][J] * A["rs"][J] if iu != 2 else A["r"][J + 1] - A["r"][J - 1],
         A["dz"][K] if iu != 3 else A["z"][K + 1] - A["z"][K - 1]]
    vol = d[0] * d[1] * d[2]
    np.testing.assert_allclose(s["bs"][I, J, K], slope_u * slope_g * vol, rtol=1e-11)


# ---------------------------------------------------------------------
# wall convective flux
# ---------------------------------------------------------------------
KAPPA_CMU = 0.4 * 0.09 ** 0.25  # von Karman constant times C_mu^(1/4)


def _qconv_cell(c, k_si, yf=1.0):
    """An open cell next to the floor of the uniform furnace with pure fuel
    (c_p of CH4, 2156 J/kg K) and turbulence energy k_si (m2/s2)."""
    A = c.arrays
    I, J, K = 8, 8, 4
    i, j, k = I - 1, J - 1, K - 1
    A["gf"][i, j, k, :] = 0.0
    A["gf"][i, j, k, IYF - 1] = yf
    A["gf"][i, j, k, IK - 1] = k_si / c.ctrl.tk0
    A["t"][i, j, k] = 1500.0 / c.ctrl.t0
    A["dnst"][i, j, k] = 0.13
    return I, J, K


def test_qconv_below_k_threshold_is_h_g():
    c = _uniform()
    I, J, K = _qconv_cell(c, 0.05)
    q = oracle.probe(c, "q_conv", I, J, K, io=[0.02, 1300.0])
    assert q == pytest.approx(c.ctrl.h_g * 200.0, rel=1e-15)


def test_qconv_cap():
    c = _uniform()
    I, J, K = _qconv_cell(c, 400.0)
    c.arrays["dnst"][I - 1, J - 1, K - 1] = 50.0
    q = oracle.probe(c, "q_conv", I, J, K, io=[0.02, 1300.0])
    assert q == pytest.approx(50.0 * 200.0, rel=1e-15)


def test_qconv_log_law_slope():
    """h_t = c_p kappa C_mu^1/4 rho sqrt(k) / (ln(E C_mu^1/4 ds rho sqrt(k)/mu) + c),
    so 1/h_t is linear in ln(ds) with slope 1/(c_p kappa C_mu^1/4 rho sqrt(k))."""
    c = _uniform()
    k_si = 4.0
    I, J, K = _qconv_cell(c, k_si)
    dT = 200.0
    h = []
    for ds in (0.01, 0.04):
        q = oracle.probe(c, "q_conv", I, J, K, io=[ds, 1500.0 - dT])
        h.append(q / dT)
    assert c.ctrl.h_g < min(h) and max(h) < 50.0
    rho = 0.13 * c.ctrl.dnst0
    slope = 1.0 / (2156.0 * KAPPA_CMU * rho * math.sqrt(k_si))
    assert (1.0 / h[1] - 1.0 / h[0]) == pytest.approx(slope * math.log(4.0), rel=1e-12)
    # sign: heat flows into the colder side
    q_neg = oracle.probe(c, "q_conv", I, J, K, io=[0.01, 1700.0])
    assert q_neg < 0.0


# ---------------------------------------------------------------------
# species sources
# ---------------------------------------------------------------------
ATOMIC = {"C": 12.011, "H": 1.008, "O": 15.999, "N": 14.007}


def test_species_sources_stoichiometry():
    c = _uniform()
    oracle.initsv_tail(c)
    oracle.sphase(c, 6)
    sl = _cell_slices(c)
    opn = c.arrays["ibcell"][sl] <= 0
    oracle.probe(c, "source", 6)
    s = oracle.scratch(c)
    crate = c.arrays["crate"][sl].copy()
    yf = c.arrays["gf"][..., IYF - 1][sl]
    assert np.count_nonzero(crate[opn] > 0.0) > 10
    fuel = (s["sfp"][sl] * yf)[opn]
    np.testing.assert_allclose(fuel, -crate[opn], rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(c.arrays["hrate"][sl][opn], crate[opn] * c.ctrl.q0, rtol=1e-15)
    src = {}
    for nelg, name in ((7, "co2"), (12, "o2"), (11, "n2")):
        oracle.probe(c, "source", nelg)
        s = oracle.scratch(c)
        src[name] = s["bs"][sl][opn]
        assert np.all(s["sfp"][sl][opn] == 0.0)
    np.testing.assert_allclose(src["co2"], crate[opn] * cs.WCO2 / cs.WF, rtol=1e-15)
    np.testing.assert_allclose(src["o2"], -crate[opn] * 2.0 * cs.WO2 / cs.WF, rtol=1e-15)
    assert np.all(src["n2"] == 0.0)
    # CH4 + 2 O2 -> CO2 + 2 H2O closes in mass with the molecular weights of
    # the atomic weights, and the code's weights agree with them
    a = ATOMIC
    w_f, w_o2 = a["C"] + 4 * a["H"], 2 * a["O"]
    w_co2, w_h2o = a["C"] + 2 * a["O"], 2 * a["H"] + a["O"]
    for w, ref in ((cs.WF, w_f), (cs.WO2, w_o2), (cs.WCO2, w_co2), (cs.WH2O, w_h2o)):
        assert w == pytest.approx(ref, rel=2e-4)
    h2o = crate[opn] * 2.0 * cs.WH2O / cs.WF
    total = fuel + src["o2"] + src["co2"] + h2o
    assert np.abs(total).max() <= 1e-3 * crate[opn].max()


# ---------------------------------------------------------------------
# CONC, DNST_CALC, THER
# ---------------------------------------------------------------------
R_U = 8314.3  # J/(kmol K), the universal gas constant


def _rho_ideal_gas(c, w):
    """rho in kg/m3 = P W / (R_u T) with P = PG0 (1 + p), T = T0 t."""
    A = c.arrays
    y = {q: A["gf"][..., q - 1][w] for q in (IYF, IYO2, IYCO2, IYN2, IYH2O)}
    w_mix = 1.0 / (y[IYF] / 16.0426 + y[IYO2] / 31.9988 + y[IYCO2] / 44.0098 + y[IYN2] / 28.0134 + y[IYH2O] / 18.0152)
    return 101325.0 * (1.0 + A["p"][w]) * w_mix / (R_U * A["t"][w] * c.ctrl.t0)


def test_conc_dnst_ther_closed_forms():
    c = _uniform()
    A = c.arrays
    I, J, K = 10, 8, 6
    i, j, k = I - 1, J - 1, K - 1
    A["gf"][i, j, k, :] = 0.0
    for q, v in ((IYF, 0.02), (IYO2, 0.1), (IYCO2, 0.12), (IYN2, 0.7)):
        A["gf"][i, j, k, q - 1] = v
    A["p"][i, j, k] = 0.01
    A["t"][i, j, k] = 1650.0 / c.ctrl.t0
    oracle.probe(c, "conc", I, J, K)
    assert A["gf"][i, j, k, IYH2O - 1] == pytest.approx(0.06, abs=1e-15)
    oracle.probe(c, "dnst_calc", I, J, K)
    w = np.zeros(c.shape, dtype=bool)
    w[i, j, k] = True
    assert A["dnst"][i, j, k] * c.ctrl.dnst0 == pytest.approx(_rho_ideal_gas(c, w)[0], rel=1e-12)
    # mass fractions above one are scaled back to sum one
    for q, v in ((IYF, 0.5), (IYO2, 0.4), (IYCO2, 0.2), (IYN2, 0.3), (IYH2O, 0.1)):
        A["gf"][i, j, k, q - 1] = v
    oracle.probe(c, "conc", I, J, K)
    ys = [A["gf"][i, j, k, q - 1] for q in (IYF, IYO2, IYCO2, IYN2, IYH2O)]
    assert sum(ys) == pytest.approx(1.0, abs=1e-15)
    assert ys[0] / ys[1] == pytest.approx(0.5 / 0.4, rel=1e-14)
    # THER: mu_t = C_mu rho k^2/eps (SI) over the reference viscosity of Re
    k_si, e_si = 3.0, 40.0
    A["gf"][i, j, k, IK - 1] = k_si / c.ctrl.tk0
    A["gf"][i, j, k, IEPS - 1] = e_si / c.ctrl.epsi0
    A["dnst"][i, j, k] = 0.2
    oracle.probe(c, "ther", I, J, K)
    rho_si = 0.2 * c.ctrl.dnst0
    mu_ref = 2.0 * c.ctrl.dnst0 * c.ctrl.ug0 * c.ctrl.r0 / c.ctrl.reyg
    assert A["tmu"][i, j, k] == pytest.approx(0.09 * rho_si * k_si ** 2 / e_si / mu_ref, rel=1e-12)
    assert A["gama"][i, j, k] == pytest.approx(2.0 / c.ctrl.reyg * (A["tmu"][i, j, k] + 1.0), rel=1e-15)


def test_enthalpy_source_history_relaxation():
    """From the second iteration on the enthalpy source is relaxed with its
    history, b = rf_h b_raw + (1 - rf_h) b_prev, and stored as the next
    history (C/Source.f90:511-520): with the history set to a known field,
    the relaxed source is that blend of the raw one (the first iteration's,
    which is not relaxed) and the field."""
    c = _uniform()
    oracle.initsv_tail(c)
    oracle.sphase(c, 3)
    sl = _cell_slices(c)
    opn = c.arrays["ibcell"][sl] <= 0
    c.ctrl.itr_gas = 1
    oracle.probe(c, "source", 8)
    raw = oracle.scratch(c)["bs"][sl].copy()
    assert np.abs(raw[opn]).max() > 0.0
    np.testing.assert_array_equal(c.arrays["bs_hprev"][sl][opn], raw[opn])
    rng = np.random.default_rng(11)
    hist = rng.uniform(-1.0, 1.0, raw.shape) * np.abs(raw[opn]).max()
    view = c.arrays["bs_hprev"][sl]
    view[...] = hist
    c.arrays["bs_hprev"][sl] = view
    c.ctrl.itr_gas = 5
    oracle.probe(c, "source", 8)
    got = oracle.scratch(c)["bs"][sl]
    rf = c.ctrl.rf_hsource
    assert rf == pytest.approx(0.3)
    np.testing.assert_allclose(got[opn], rf * raw[opn] + (1.0 - rf) * hist[opn], rtol=1e-13,
                               atol=1e-13 * np.abs(raw[opn]).max())
    np.testing.assert_array_equal(c.arrays["bs_hprev"][sl][opn], got[opn])


# ---------------------------------------------------------------------
# half-blocked cells (IBCELL -1, -2; C/Grid3.f90:243-256,389-471)
# ---------------------------------------------------------------------
HALF_BLOCKED = [((6, 8), (3, 5), (2, 4), -1), ((15, 15), (7, 9), (6, 7), -2), ((0, 0), (5, 5), (4, 4), -1)]
BV_F = (0.5, 0.7)
BA_F = (0.6, 0.8, 0.9, 0.75, 0.65, 0.85)


def half_blocked_case():
    return cs.furnace_small(nx=16, ny=12, nz=10, seed=3, half_blocked=HALF_BLOCKED, bv_f=BV_F, ba_f=BA_F)


def _base_area(c, I, J, K, ix):
    A = c.arrays
    dx, dr, dz, rs = A["dx"][I - 1], A["dr"][J - 1], A["dz"][K - 1], A["rs"][J - 1]
    return (dr * rs * dz, dx * dz, dx * dr * rs)[ix - 1]


def test_half_blocked_cells_are_open_and_not_propagated():
    c = half_blocked_case()
    ib = c.arrays["ibcell"]
    assert ib[2 * 6 + 3, 2 * 3 + 3, 2 * 2 + 3] == -1 and ib[2 * 15 + 3, 2 * 7 + 3, 2 * 6 + 3] == -2
    # a half-blocked cell's faces keep the open code (GRID3 propagates codes >= 1 only)
    assert ib[2 * 6 + 3 + 1, 2 * 3 + 3, 2 * 2 + 3] == 0
    plain = cs.furnace_small(nx=16, ny=12, nz=10, seed=3)
    assert c.ctrl.ncells == plain.ctrl.ncells


def test_half_blocked_area_and_volume_closed_forms():
    c = half_blocked_case()
    oracle.initsv_tail(c)
    ba = np.array(BA_F).reshape(2, 3)  # [n-1][ix-1]
    mp = c.shape[0]
    cases = []
    I, J, K = 4 + 2 * 7, 4 + 2 * 4, 4 + 2 * 3          # inside the -1 block
    for ix in (1, 2, 3):
        cases.append(((I, J, K), ix, _base_area(c, I, J, K, ix) * ba[0][ix - 1]))
    I2, J2, K2 = 4 + 2 * 15, 4 + 2 * 8, 4 + 2 * 6      # a -2 cell on the last plane
    for ix in (1, 2, 3):
        cases.append(((I2, J2, K2), ix, _base_area(c, I2, J2, K2, ix) * ba[1][ix - 1]))
    # the exit cell beyond it (code 3) takes BA_F of its half-blocked x neighbour
    assert c.arrays["ibcell"][mp - 1, J2 - 1, K2 - 1] == 3
    cases.append(((mp, J2, K2), 1, _base_area(c, mp, J2, K2, 1) * ba[1][0]))
    # the inlet cell next to the -1 cell at the inlet wall, likewise
    assert c.arrays["ibcell"][1, 4 + 2 * 5 - 1, 4 + 2 * 4 - 1] == 2
    cases.append(((2, 4 + 2 * 5, 4 + 2 * 4), 1, _base_area(c, 2, 4 + 2 * 5, 4 + 2 * 4, 1) * ba[0][0]))
    # an open cell next to the block: its own area
    cases.append(((I + 6, J, K), 1, _base_area(c, I + 6, J, K, 1)))
    for (i, j, k), ix, want in cases:
        got = oracle.probe(c, "area_c", i, j, k, io=[ix, 0.0])
        assert got == pytest.approx(want, rel=1e-15), ((i, j, k), ix)
    A = c.arrays
    v0 = lambda i, j, k: A["dx"][i - 1] * A["dr"][j - 1] * A["rs"][j - 1] * A["dz"][k - 1]
    assert oracle.probe(c, "vol_c", I, J, K) == pytest.approx(v0(I, J, K) * BV_F[0], rel=1e-15)
    assert oracle.probe(c, "vol_c", I2, J2, K2) == pytest.approx(v0(I2, J2, K2) * BV_F[1], rel=1e-15)
    # x face between the last -1 cell of the block and the open cell beyond:
    # half of each cell's volume, the half-blocked one scaled by BV_F
    Ie = 4 + 2 * 8
    want = 0.5 * v0(Ie, J, K) * BV_F[0] + 0.5 * v0(Ie + 2, J, K)
    assert oracle.probe(c, "vol_c", Ie + 1, J, K) == pytest.approx(want, rel=1e-15)


def test_half_blocked_staggered_volume_of_the_buoyancy_source():
    """VOL1 (C/Gslv3x.f90:266-286): a w face next to a half-blocked cell
    loses half of that cell's height times the face area, so the buoyancy
    source there is rho g ((Z(k+1) - Z(k-1)) - DZ(k +- 1)/2) A."""
    c = half_blocked_case()
    A = c.arrays
    A["theta"][...] = 1.0
    A["dnst"][...] = 1.0
    A["gama"][...] = 0.02
    A["ug"][...] = 0.0
    g = c.ctrl.gra[2]
    oracle.probe(c, "momentum")
    bs = oracle.scratch(c)["bs"]
    I, J = 4 + 2 * 7, 4 + 2 * 4
    z, dz = A["z"], A["dz"]
    for K, below, above in ((4 + 2 * 2 - 1, False, True), (4 + 2 * 3 + 1, True, True), (4 + 2 * 4 + 1, True, False)):
        area = _base_area(c, I, J, K, 3)
        vol = (z[K] - z[K - 2]) * area  # Z(K+1) - Z(K-1), 1-based
        if below:
            vol -= dz[K - 2] * area / 2.0
        if above:
            vol -= dz[K] * area / 2.0
        assert bs[I - 1, J - 1, K - 1] == pytest.approx(g * vol, rel=1e-12), K
