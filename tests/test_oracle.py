"""CPU checks that pin the oracle (oracle/gfm_oracle.c) before it is used
to judge the GPU path.

What pins what (each check is chosen so a plausible slip -- a dropped
term, a wrong sign, a swapped index -- fails it):
  DA (C/Prop.f90:97-117)          Patankar's exponential scheme closed form
                                  A(P)=P/(e^P-1), its series below P=0.01
                                  and the cut-off above P=20.
  TGAMF (C/Gflow.f90:417-426)     log-law closed form and the 11.5 switch.
  T_to_enth / enth_to_T           pure-species closed forms including the
  (C/Prop.f90:369-565)            reference's non-accumulated F/H terms;
                                  T->h->T round trip; bisection on T_to_enth.
  ZY1 (C/Source.f90:814-831)      piecewise-linear closed form.
  ADLBL3 / calc_residual          manufactured solution on a random
  (C/Adlbl3.f90)                  diagonally dominant system (both plane
                                  orders converge to it; residual -> 0;
                                  blocked cells untouched); calc_residual
                                  against an independent numpy evaluation.
  SOURCE(5) res_mass              closed form for a single perturbed face.
  FLX1D (C/Flx1d.f90:16-73)       an independent numpy evaluation of the
                                  per-plane flow sums on a developed state.
  x-plane block correction        its defining property: afterwards the
  (no reference counterpart)      residuals of each plane's unpinned cells
                                  sum to zero (numpy evaluation) on every
                                  plane but the last (the reference plane,
                                  left unchanged), every such cell moved by
                                  its plane's one value, and pinned and
                                  out-of-box cells unchanged.
  case generator                  closed-form lattice coordinates, brute-force
                                  cell-code propagation on a tiny grid,
                                  inflow = GMFR, nondimensional groups.
The composite GFLOW has no closed form; it is pinned only through these
parts ("parity unpinned" as a whole against the reference Fortran, which
cannot be built here).
"""
import ctypes as ct
import math

import numpy as np
import pytest

import oracle
from gfm_amd import case as cs
from gfm_amd.abi import host_arrays
from helpers import manufactured


# ---------------- DA / TGAMF ----------------
@pytest.mark.parametrize("pe", [1e-4, 5e-3, 1e-2, 0.02, 0.5, 3.0, 19.9])
def test_da_closed_form(pe):
    d = 2.5
    f = pe * d
    exact = d * pe / math.expm1(pe)  # Patankar A(|P|) = |P|/(exp|P|-1)
    assert oracle.da(d, f) == pytest.approx(exact, rel=1e-9)
    assert oracle.da(d, -f) == pytest.approx(exact, rel=1e-9)


def test_da_cutoffs():
    assert oracle.da(1.0, 20.0001) == 0.0
    assert oracle.da(1e-26, 1.0) == 0.0
    assert oracle.da(1.0, 0.0) == 1.0


def test_tgamf():
    assert oracle.tgamf(11.49) == 1.0
    for b in (11.5, 30.0, 1e3):
        assert oracle.tgamf(b) == pytest.approx(max(b / (2.5 * math.log(9 * b)), 1.0), rel=1e-14)


# ---------------- enthalpy <-> temperature ----------------
NIST = dict(A=cs.A_N, B=cs.B_N, C=cs.C_N, D=cs.D_N, E=cs.E_N, F=cs.F_N, H=cs.H_N)
W = [cs.WF, cs.WO2, cs.WCO2, cs.WN2, cs.WH2O]


def shomate_pure_quirk(n, T):
    """h of pure species n (mass fraction one) as the reference forms it:
    the F and H terms are taken from the running D and E sums plus the
    last species' (H2O) F and H weighted by its mass fraction, so for a
    pure species other than H2O they become D_n and E_n."""
    s = 0 if T < cs.CUT_T[n] else 1
    t = T / 1000.0
    a, b, c, d, e = (NIST[k][n][s] for k in "ABCDE")
    if n == 4:
        f, h = d + NIST["F"][n][s], e + NIST["H"][n][s]
    else:
        f, h = d, e
    g = t * (a + t * (b / 2 + t * (c / 3 + t * d / 4))) - e / t + f - h
    return g * 1e6 / (W[n] * cs.H_0)


@pytest.mark.parametrize("n", range(5))
@pytest.mark.parametrize("T", [350.0, 1150.0, 1250.0, 1750.0, 2400.0])
def test_t_to_enth_pure_species(n, T):
    y = np.zeros(5)
    y[n] = 1.0
    assert oracle.t_to_enth(y, T / cs.T0, cs.H_0, cs.T0) == pytest.approx(shomate_pure_quirk(n, T), rel=1e-12)


def test_t_to_enth_matches_python_case_generator():
    rng = np.random.default_rng(0)
    for _ in range(50):
        y = rng.dirichlet(np.ones(5))
        T = rng.uniform(250, 2900)
        assert oracle.t_to_enth(y, T / cs.T0, cs.H_0, cs.T0) == pytest.approx(
            float(cs.t_to_enth(y[:, None], np.array([T / cs.T0]))[0]), rel=1e-13)


def test_enth_to_t_round_trip_and_bisection():
    """Without a lo/hi coefficient switch inside Newton the inversion is
    exact to the tolerance.  (When Newton crosses a cut temperature the
    reference rebuilds f_mix/h_mix from the switched species instead of
    H2O, C/Prop.f90:465-494, which moves the root; the oracle keeps that.)"""
    rng = np.random.default_rng(1)
    cuts = cs.CUT_T[[0, 2, 4]]
    n = 0
    while n < 200:
        y = rng.dirichlet(np.ones(5) * 0.7)
        T = rng.uniform(260, 2950)
        guess = T * rng.uniform(0.95, 1.05)
        if np.any(np.abs(T - cuts) < 0.08 * T):
            continue
        n += 1
        h = oracle.t_to_enth(y, T / cs.T0, cs.H_0, cs.T0)
        Tn = oracle.enth_to_t(y, h, guess / cs.T0, cs.H_0, cs.T0) * cs.T0
        lo, hi = 200.0, 3200.0
        for _ in range(80):
            mid = 0.5 * (lo + hi)
            if oracle.t_to_enth(y, mid / cs.T0, cs.H_0, cs.T0) < h:
                lo = mid
            else:
                hi = mid
        assert abs(Tn - T) < 1e-6
        assert abs(0.5 * (lo + hi) - T) < 1e-6


# ---------------- ZY1 ----------------
def test_zy1():
    x = np.array([1.0, 2.0, 4.0, 8.0])
    y = np.array([10.0, 20.0, 25.0, 40.0])
    assert oracle.zy1(x, y, 0.5) == pytest.approx(5.0)
    assert oracle.zy1(x, y, 3.0) == pytest.approx(22.5)
    assert oracle.zy1(x, y, 2.0) == pytest.approx(20.0)
    assert oracle.zy1(x, y, 9.0) == 40.0


# ---------------- ADLBL3 / calc_residual ----------------
@pytest.mark.parametrize("order", [0, 1])
def test_adlbl3_manufactured_solution(order):
    c = cs.furnace_small(nx=10, ny=8, nz=8, seed=5)
    oracle.initsv_tail(c)
    box = (4, c.ctrl.mp - 2, 4, c.ctrl.np - 2, 4, c.ctrl.lp - 2)
    ap, as_, bs, xt, xs, sl = manufactured(c, box, seed=11)
    c.ctrl.sweep_order = order
    c.ctrl.ntimesg[3] = 60
    r0 = oracle.calc_residual(c, ap, as_, bs, xt, box)
    oracle.adlbl3(c, 4, ap, as_, bs, xt, box)
    r1 = oracle.calc_residual(c, ap, as_, bs, xt, box)
    ibc = c.arrays["ibcell"][sl]
    opn = ibc <= 0
    assert np.max(np.abs(xt[sl][opn] - xs[sl][opn])) < 1e-10
    assert np.all(xt[sl][~opn] == 0.0)  # blocked rows keep their value
    assert r1 < 1e-10 * r0


def _np_residual(c, ap, as_, bs, xt, box):
    """Residual sum a_nb x_nb + b - a_p x on the box's cells (numpy)."""
    ib, ie, jb, je, kb, ke = box
    sl = (slice(ib - 1, ie, 2), slice(jb - 1, je, 2), slice(kb - 1, ke, 2))

    def sh(di, dj, dk):
        return xt[ib - 1 + di:ie + di:2, jb - 1 + dj:je + dj:2, kb - 1 + dk:ke + dk:2]
    r = (as_[..., 0, 0][sl] * sh(-2, 0, 0) + as_[..., 0, 1][sl] * sh(2, 0, 0) + as_[..., 1, 0][sl] * sh(0, -2, 0)
         + as_[..., 1, 1][sl] * sh(0, 2, 0) + as_[..., 2, 0][sl] * sh(0, 0, -2) + as_[..., 2, 1][sl] * sh(0, 0, 2)
         + bs[sl] - ap[sl] * xt[sl])
    return r, sl


@pytest.mark.parametrize("seed", [1, 2])
def test_block_corr_x_zeroes_plane_residual_sums(seed):
    c = cs.furnace_small(nx=12, ny=8, nz=6, seed=seed)
    oracle.initsv_tail(c)
    box = (4, c.ctrl.mp - 2, 4, c.ctrl.np - 2, 4, c.ctrl.lp - 2)
    ap, as_, bs, xt, xs, sl = manufactured(c, box, seed=20 + seed)
    rng = np.random.default_rng(seed)
    xt[sl] = rng.uniform(-1, 1, xt[sl].shape)
    # pin a few open interior rows through |AP| <= 1e-20 as well
    ibc = c.arrays["ibcell"]
    opn = np.argwhere(ibc[sl] <= 0)
    for q in opn[rng.choice(len(opn), 5, replace=False)]:
        ap[tuple(2 * q + np.array([box[0] - 1, box[2] - 1, box[4] - 1]))] = 0.0
    x0 = xt.copy(order="F")
    r0, _ = _np_residual(c, ap, as_, bs, x0, box)
    free = (ibc[sl] <= 0) & (np.abs(ap[sl]) > 1e-20)
    oracle.block_corr_x(c, ap, as_, bs, xt, box)
    r1, _ = _np_residual(c, ap, as_, bs, xt, box)
    d = xt - x0
    outside = np.ones(xt.shape, bool)
    outside[sl] = False
    assert np.all(d[outside] == 0.0)
    assert np.all(d[sl][~free] == 0.0)
    scale = np.abs(r0).max()
    planes = 0
    last = free.shape[0] - 1
    assert np.all(d[sl][last] == 0.0)  # the reference plane
    for q in range(last):
        f = free[q]
        if not f.any():
            continue
        planes += 1
        dq = d[sl][q][f]
        assert np.ptp(dq) <= 1e-14 * max(1.0, np.abs(dq).max()), q  # one value per plane
        assert abs(r1[q][f].sum()) <= 1e-11 * scale * f.sum(), q
        assert abs(r0[q][f].sum()) > 1e-3 * scale  # there was something to correct
        assert np.abs(dq).max() > 0.0
    assert planes >= 8


def test_block_corr_x_nearly_singular_system_stays_bounded():
    """A p'-like system: no coefficient toward pinned or out-of-box cells
    and a_p = sum a_nb up to FP32 rounding (relative 1e-7), so the plane
    system is singular but for rounding.  The correction stays of the size
    of the solution (solving the full system would divide the total residual
    by the rounding-level last pivot) and still zeroes the plane sums."""
    c = cs.furnace_small(nx=12, ny=8, nz=6, seed=4)
    oracle.initsv_tail(c)
    box = (4, c.ctrl.mp - 2, 4, c.ctrl.np - 2, 4, c.ctrl.lp - 2)
    ap, as_, bs, xt, xs, sl = manufactured(c, box, seed=30)
    ibc = c.arrays["ibcell"]
    ib, ie, jb, je, kb, ke = box
    free = np.zeros(ibc.shape, bool)
    free[sl] = ibc[sl] <= 0
    fp = np.pad(free, 2)
    shp = ibc.shape
    for d, (di, dj, dk) in enumerate([(-2, 0, 0), (0, -2, 0), (0, 0, -2), (2, 0, 0), (0, 2, 0), (0, 0, 2)]):
        nbf = fp[2 + di:2 + di + shp[0], 2 + dj:2 + dj + shp[1], 2 + dk:2 + dk + shp[2]]
        as_[..., d % 3, d // 3][~nbf] = 0.0
    rng = np.random.default_rng(5)
    ap[sl] = as_[sl].sum(axis=(-1, -2)) * (1.0 + 1e-7 * rng.uniform(-1, 1, ap[sl].shape))
    xt[sl] = rng.uniform(-1, 1, xt[sl].shape)
    x0 = xt.copy(order="F")
    oracle.block_corr_x(c, ap, as_, bs, xt, box)
    d = (xt - x0)[sl]
    r1, _ = _np_residual(c, ap, as_, bs, xt, box)
    fr = free[sl]
    print("max |d|", np.abs(d).max(), flush=True)
    assert 0.0 < np.abs(d).max() < 10.0 * (np.abs(xs).max() + 1.0)
    scale = np.abs(_np_residual(c, ap, as_, bs, x0, box)[0]).max()
    for q in range(fr.shape[0] - 1):
        if fr[q].any():
            assert abs(r1[q][fr[q]].sum()) <= 1e-9 * scale * fr[q].sum(), q


def test_flx1d_matches_numpy():
    c = cs.furnace_small(nx=10, ny=8, nz=6, seed=3)
    oracle.initsv_tail(c)
    oracle.sphase(c, 3)
    oracle.flx1d(c)
    got = oracle.plane_flows(c)
    a, k = c.arrays, c.ctrl
    mp, np_, lp = c.shape
    ib = a["ibcell"].astype(int)
    skip = (ib == 1) | (ib >= 4)
    dx, dr, rs, dz = a["dx"], a["dr"], a["rs"], a["dz"]
    area = [dr[None, :, None] * rs[None, :, None] * dz[None, None, :],
            dx[:, None, None] * dz[None, None, :],
            dx[:, None, None] * dr[None, :, None] * rs[None, :, None]]
    g = [np.where(skip | (ib == 1), 0.0, a["theta"] * a["dnst"] * a["ug"][..., d] * area[d] * k.gmfr) for d in range(3)]
    h = a["gf"][..., 2] * k.h_0
    # x-faces: I = 3..MP-1 over J = 2..NP, K = 2..LP (1-based)
    for I in range(3, mp, 2):
        v = g[0][I - 1, 1::2, 1::2]
        want = (v[v >= 0].sum(), -v[v < 0].sum())
        np.testing.assert_allclose(got["flx"][I - 1], want, rtol=1e-12, atol=1e-300)
    for J in range(3, np_, 2):
        v = g[1][3:mp - 2:2, J - 1, 3:lp - 2:2]
        hv = v * h[3:mp - 2:2, J - 1, 3:lp - 2:2]
        want = (v[v >= 0].sum(), -v[v < 0].sum())
        np.testing.assert_allclose(got["fly"][J - 1], want, rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(got["flyh"][J - 1], (hv[v >= 0].sum(), -hv[v < 0].sum()), rtol=1e-12, atol=1e-300)
    for K in range(3, lp, 2):
        v = g[2][3:mp - 2:2, 3:np_ - 2:2, K - 1]
        hv = v * h[3:mp - 2:2, 3:np_ - 2:2, K - 1]
        want = (v[v >= 0].sum(), -v[v < 0].sum())
        np.testing.assert_allclose(got["flux_z"][K - 1], want, rtol=1e-12, atol=1e-300)
        np.testing.assert_allclose(got["fluxh_z"][K - 1], (hv[v >= 0].sum(), -hv[v < 0].sum()), rtol=1e-12, atol=1e-300)
    # even indices are not planes
    assert np.all(got["flx"][1::2] == 0.0) and np.all(got["fly"][1::2] == 0.0)
    # the flow enters through the inlet (x = 0 wall) and leaves through the exit face
    assert got["flx"][2, 0] > 0.0 and got["flx"][mp - 2, 0] > 0.0


def test_calc_residual_definition():
    c = cs.furnace_small(nx=8, ny=8, nz=8, seed=6)
    box = (4, c.ctrl.mp - 2, 4, c.ctrl.np - 2, 4, c.ctrl.lp - 2)
    rng = np.random.default_rng(4)
    shp = c.shape
    ap = np.asfortranarray(rng.uniform(1, 2, shp))
    as_ = np.asfortranarray(rng.uniform(0, 1, shp + (3, 2)))
    bs = np.asfortranarray(rng.uniform(-1, 1, shp))
    xt = np.asfortranarray(rng.uniform(-1, 1, shp))
    got = oracle.calc_residual(c, ap, as_, bs, xt, box)
    ib = c.arrays["ibcell"]
    tot = 0.0
    for i in range(box[0], box[1] + 1, 2):
        for j in range(box[2], box[3] + 1, 2):
            for k in range(box[4], box[5] + 1, 2):
                I, J, K = i - 1, j - 1, k - 1
                if ib[I, J, K] > 0:
                    continue
                terms = [as_[I, J, K, 0, 0] * xt[I - 2, J, K], as_[I, J, K, 0, 1] * xt[I + 2, J, K],
                         as_[I, J, K, 1, 0] * xt[I, J - 2, K], as_[I, J, K, 1, 1] * xt[I, J + 2, K],
                         as_[I, J, K, 2, 0] * xt[I, J, K - 2], as_[I, J, K, 2, 1] * xt[I, J, K + 2]]
                r = sum(terms) + bs[I, J, K] - ap[I, J, K] * xt[I, J, K]
                tot += abs(r) / max([abs(t) for t in terms] + [abs(ap[I, J, K] * xt[I, J, K]), abs(bs[I, J, K])])
    assert got == pytest.approx(max(tot / c.ncells, 1e-16), rel=1e-12)


# ---------------- res_mass ----------------
def test_res_mass_single_face():
    c = cs.furnace_small(nx=10, ny=8, nz=8, seed=7)
    oracle.initsv_tail(c)
    c.arrays["ug"][...] = 0.0
    I, J, K = 11, 8, 8  # interior x-face between centres 10 and 12
    c.arrays["ug"][I - 1, J - 1, K - 1, 0] = 0.3
    h = host_arrays(c.arrays)
    lib = oracle.lib()
    lib.oracle_mass_residual.restype = ct.c_double
    rm = lib.oracle_mass_residual(ct.byref(c.ctrl), ct.byref(h))
    assert rm == pytest.approx(2.0 / c.ncells, rel=1e-14)
    assert c.ctrl.res_mass_max == pytest.approx(1.0, rel=1e-14)
    assert (c.ctrl.imx, c.ctrl.jmx, c.ctrl.kmx) == (10, 8, 8)


# ---------------- case generator ----------------
def test_lattice_coordinates_uniform():
    x, dx = cs.lattice_coord(np.linspace(0, 3.0, 49))
    assert len(x) == 100
    assert x[0] == x[1] == x[2] == 0.0 and x[-1] == x[-2] == 3.0
    for c in range(1, 49):  # centre of cell c-1 at index I=2c+2 (1-based)
        assert x[2 * c + 1] == pytest.approx((c - 0.5) * 3.0 / 48)
    assert dx[3] == pytest.approx(3.0 / 48)     # centre I=4: cell width
    assert dx[4] == pytest.approx(3.0 / 48)     # face I=5: centre spacing
    assert dx[0] == 0.0 and dx[-1] == 0.0


def test_ibcell_propagation_brute_force():
    rng = np.random.default_rng(9)
    mp, np_, lp = 12, 10, 10
    ib = np.zeros((mp, np_, lp), dtype=np.int16)
    cen = ib[1::2, 1::2, 1::2]
    cen[...] = rng.choice([0, 0, 0, 1, 2, 3], size=cen.shape)
    ref = ib.copy()
    ncells = cs.propagate_ibcell(ib)
    assert ncells == int(np.count_nonzero(cen <= 0))
    # brute force: every non-centre position takes 1 if any adjacent centre is a
    # wall, else the code of the last adjacent non-open centre in I,J,K order
    for i in range(mp):
        for j in range(np_):
            for k in range(lp):
                if i % 2 == 1 and j % 2 == 1 and k % 2 == 1:
                    continue
                codes = []
                for ci in range(max(i - 1, 0), min(i + 2, mp)):
                    for cj in range(max(j - 1, 0), min(j + 2, np_)):
                        for ck in range(max(k - 1, 0), min(k + 2, lp)):
                            if ci % 2 == 1 and cj % 2 == 1 and ck % 2 == 1 and ref[ci, cj, ck] >= 1:
                                codes.append(int(ref[ci, cj, ck]))
                want = 1 if 1 in codes else (codes[-1] if codes else 0)
                assert ib[i, j, k] == want, (i, j, k)


def test_inflow_equals_gmfr_and_groups():
    c = cs.furnace_small()
    assert c.ctrl.gfin0 == pytest.approx(1.0, rel=1e-12)
    ug0 = c.ctrl.ug0
    dnst0 = cs.PG0 * cs.WT0 / cs.RU / cs.T0
    assert c.ctrl.dnst0 == pytest.approx(dnst0)
    assert c.ctrl.reyg == pytest.approx(2 * dnst0 * ug0 / cs.GMU_REF)
    assert c.ctrl.eun == pytest.approx(cs.PG0 / (dnst0 * ug0 ** 2))
    # fuel 2x2 cells of 0.05 m at 30 m/s, air 20 cells at 10 m/s
    rho_f = cs.PG0 * cs.WF / (cs.RU * 300.0)
    assert c.meta["fr_f"] == pytest.approx(rho_f * 30.0 * 4 * 0.05 * 0.05, rel=1e-12)
    assert c.ncells == 48 * 24 * 16


def test_kinetic_tables():
    k = cs.load_kinetic()
    assert all(len(v) == 100 for v in k.values())
    assert np.all(np.diff(k["rtim"]) > 0) and np.all(np.diff(k["delr"]) > 0)
    assert k["rtim"][0] == pytest.approx(4.5e-7 * 0.7)
    assert k["delr"][-1] == 1.0


def test_oracle_iterates_and_decreases_residual():
    c = cs.furnace_small(nx=16, ny=12, nz=10, seed=3)
    oracle.initsv_tail(c)
    h = oracle.sphase(c, 40)
    assert np.all(np.isfinite(h))
    assert h[-1] < 0.25 * h[0]


def test_emf_closed_form():
    lib = oracle.lib()
    lib.oracle_emf.restype = ct.c_double
    lib.oracle_emf.argtypes = [ct.c_double, ct.c_double]
    for x in (0.5, 2.0, 0.002):
        assert lib.oracle_emf(x, 1.0) == pytest.approx(-math.expm1(-x), rel=1e-14)
    g1 = -math.expm1(-1e-3) / 1e-3
    assert lib.oracle_emf(1e-4, 2.0) == pytest.approx(2e-4 * g1, rel=1e-14)
    # optically thin limit of the gray emission: SA*EMF(kappa, 6V/SA) -> 6 kappa V
    assert lib.oracle_emf(0.3, 1e-3) * 6.0 / 1e-3 == pytest.approx(6 * 0.3 * g1, rel=1e-14)


@pytest.mark.parametrize("walls,kappa", [("isothermal", 0.3), ("adiabatic", 0.0)])
def test_physics_variants_iterate(walls, kappa):
    c = cs.furnace_small(nx=12, ny=10, nz=8, seed=8, walls=walls, gray_kappa=kappa)
    oracle.initsv_tail(c)
    h = oracle.sphase(c, 10)
    assert np.all(np.isfinite(h))
    if walls == "isothermal":
        assert c.ctrl.q_wall != 0.0
    else:
        assert c.ctrl.q_wall == 0.0
