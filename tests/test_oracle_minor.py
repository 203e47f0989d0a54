"""Pins of the oracle's minor-species solve (MSFLOW C/Msflow.f90, GSLV3M
C/Gslv3m.f90, SOURCEM C/Sourcem.f90, INITSVM C/Initsv.f90:329-380, extrm
C/Extr.f90:258-305, intpm C/Intp.f90:348-419): each is compared with an
invariant or a law that does not restate its formula.

  SOURCEM species table            per cell the seven sources conserve
                                   mass and the C, H, O and N atoms.
  GSLV3M rates                     each rate scales with its reactant
                                   concentrations by its reaction orders
                                   and with T by exp(-E/RT) (T^beta for
                                   NO), independently per reactant.
  RR2 carried between cells        a cell without CH4 whose CO oxidation
                                   is limited by the O2 inflow loses 0.9
                                   times the NO-rate change of the
                                   previous open cell; a cell with CH4
                                   does not.
  GSLV3M transport                 a uniform field is a fixed point of
                                   every species and the soot equation;
                                   from mass fractions summing to one,
                                   one GSLV3M keeps the sum (bal_ms at
                                   rounding) while no clamp binds.
  soot sources                     formation linear in CH4 and zero below
                                   T0_sf, oxidation a sink linear in O2,
                                   both relaxed against the history after
                                   the first iteration.
  intpm, extrm                     intpm is INTP's loop and weights (the
                                   same face values from the same cell
                                   values); exit cells extrapolate linear
                                   fields exactly, walls copy.
  MSFLOW                           the convergence test ends at itr_ms 11
                                   before extrm/intpm; maxms bounds it.
"""
import numpy as np
import pytest

import oracle
from gfm_amd import case as cs
from gfm_amd import minor

W = np.array([16.0426, 31.9988, 28.0104, 18.0152, 44.0098, 28.0134, 30.0061])
# atoms per molecule: C, H, O, N of CH4, O2, CO, H2O, CO2, N2, NO
ATOMS = np.array([[1, 4, 0, 0], [0, 0, 2, 0], [1, 0, 1, 0], [0, 2, 1, 0], [1, 0, 2, 0], [0, 0, 0, 2], [0, 0, 1, 1]],
                 dtype=float)


def _developed(nx=24, ny=12, nz=8, iters=40):
    c = cs.furnace_small(nx=nx, ny=ny, nz=nz, seed=3)
    oracle.initsv_tail(c)
    oracle.sphase(c, iters)
    return c


@pytest.fixture(scope="module")
def developed():
    return _developed()


def _cen(c):
    ib = c.arrays["ibcell"]
    m = np.zeros(ib.shape, dtype=bool)
    m[3:-2:2, 3:-2:2, 3:-2:2] = ib[3:-2:2, 3:-2:2, 3:-2:2] <= 0
    return m


def _with_co_no(c, ms, rng):
    """INITSVM's fresh state plus CO and NO fields, renormalised, so that
    every reaction runs."""
    oracle.initsvm(c, ms)
    g = ms.arrays["gfm"]
    cen = (slice(1, None, 2),) * 3
    g[cen + (2,)] = rng.uniform(0.0, 0.03, g[cen + (2,)].shape)
    g[cen + (6,)] = rng.uniform(0.0, 1e-3, g[cen + (6,)].shape)
    s = g[cen].sum(axis=-1, keepdims=True)
    g[cen] = g[cen] / np.where(s > 0, s, 1.0)


def test_species_sources_conserve_mass_and_atoms(developed):
    c = developed
    ms = minor.MinorSpecies(c)
    _with_co_no(c, ms, np.random.default_rng(1))
    oracle.ms_probe(c, ms, "rates")
    rr = ms.arrays["rrms"]
    assert rr[..., 0].max() > 0 and rr[..., 1].max() > 0 and rr[..., 3].max() > 0
    cen = _cen(c)
    bs = []
    for l in range(1, 8):
        oracle.ms_probe(c, ms, "sourcem", l)
        bs.append(oracle.scratch(c)["bs"][cen])
    bs = np.array(bs)  # (7, cells), kg/s / GMFR
    scale = np.abs(bs).sum(axis=0) + 1e-300
    assert np.abs(bs.sum(axis=0) / scale).max() < 1e-12
    mol = bs / W[:, None]
    for e in range(4):
        bal = (mol * ATOMS[:, e:e + 1]).sum(axis=0)
        assert np.abs(bal / (np.abs(mol).sum(axis=0) + 1e-300)).max() < 1e-12, "CHON"[e]


def _one_cell(c, ms, q, setup):
    """The rates with every rate history zero, after setup(gfm, t) at cell q."""
    g = ms.arrays["gfm"]
    t = c.arrays["t"]
    g0, t0 = g.copy(), t.copy()
    setup(g, t)
    ms.arrays["rrms"][...] = 0.0
    oracle.ms_probe(c, ms, "rates")
    out = ms.arrays["rrms"][q[0] // 2 - 1, q[1] // 2 - 1, q[2] // 2 - 1, :].copy()
    g[...] = g0
    t[...] = t0
    return out


def test_rate_laws_follow_reaction_orders_and_arrhenius(developed):
    """Reaction 1: CH4^1.46 O2^0.5217 exp(-20643/T); 2: CO^1.6904 O2^1.57
    exp(-11613/T); 5: N2 O2^0.5 T^-0.5 exp(-68899/T); 4: CO^0.721
    O2^4.0111 Y_N2 exp(-44369/T) (Var_mod ODR, ERC, BETA).  At a cool cell
    no inflow limit binds and each rate scales by the ratios alone."""
    c = developed
    ms = minor.MinorSpecies(c)
    oracle.initsvm(c, ms)
    q = (c.shape[0] // 2 // 2 * 2, c.shape[1] // 2 // 2 * 2, c.shape[2] // 2 // 2 * 2)
    assert c.arrays["ibcell"][q[0] - 1, q[1] - 1, q[2] - 1] == 0
    t0 = c.ctrl.t0
    I = (q[0] - 1, q[1] - 1, q[2] - 1)

    nbs = [(I[0] + a, I[1] + b, I[2] + d) for a, b, d in
           ((-2, 0, 0), (2, 0, 0), (0, -2, 0), (0, 2, 0), (0, 0, -2), (0, 0, 2))]

    def state(y1, y2, y3, y6, tk, y5=0.0):
        def f(g, t):
            for n in nbs:  # ample inflow of every reactant: no inflow limit binds
                g[n + (slice(None),)] = 0.0
                g[n + (0,)], g[n + (1,)], g[n + (2,)], g[n + (4,)], g[n + (5,)] = 0.2, 0.2, 0.1, 0.1, 0.4
            g[I + (slice(None),)] = 0.0
            g[I + (0,)], g[I + (1,)], g[I + (2,)], g[I + (5,)], g[I + (4,)] = y1, y2, y3, y6, y5
            t[I] = tk / t0
        return f

    base = _one_cell(c, ms, q, state(0.02, 0.1, 0.0, 0.7, 700.0))
    assert base[0] > 0
    r = _one_cell(c, ms, q, state(0.04, 0.1, 0.0, 0.7, 700.0))
    assert r[0] / base[0] == pytest.approx(2 ** 1.46, rel=1e-12)
    r = _one_cell(c, ms, q, state(0.02, 0.2, 0.0, 0.7, 700.0))
    assert r[0] / base[0] == pytest.approx(2 ** 0.5217, rel=1e-12)
    r = _one_cell(c, ms, q, state(0.02, 0.1, 0.0, 0.7, 720.0))
    assert r[0] / base[0] == pytest.approx(np.exp(-20643.0 / 720 + 20643.0 / 700), rel=1e-12)
    # CO oxidation, without CH4 (reaction 1 skipped) and CO2
    b2 = _one_cell(c, ms, q, state(0.0, 0.1, 0.01, 0.7, 500.0))
    assert b2[1] > 0 and b2[0] == 0.0
    r = _one_cell(c, ms, q, state(0.0, 0.1, 0.02, 0.7, 500.0))
    assert r[1] / b2[1] == pytest.approx(2 ** 1.6904, rel=1e-12)
    r = _one_cell(c, ms, q, state(0.0, 0.2, 0.01, 0.7, 500.0))
    assert r[1] / b2[1] == pytest.approx(2 ** 1.57, rel=1e-12)
    # NO: thermal (reaction 5) alone without CO, then reaction 4 as the CO part
    n5 = _one_cell(c, ms, q, state(0.0, 0.1, 0.0, 0.7, 2000.0))[3]
    assert n5 > 0
    assert _one_cell(c, ms, q, state(0.0, 0.1, 0.0, 0.35, 2000.0))[3] / n5 == pytest.approx(0.5, rel=1e-12)
    assert _one_cell(c, ms, q, state(0.0, 0.2, 0.0, 0.7, 2000.0))[3] / n5 == pytest.approx(2 ** 0.5, rel=1e-12)
    ratio = np.exp(-68899.0 / 2100 + 68899.0 / 2000) * (2100.0 / 2000.0) ** -0.5
    assert _one_cell(c, ms, q, state(0.0, 0.1, 0.0, 0.7, 2100.0))[3] / n5 == pytest.approx(ratio, rel=1e-12)
    n45 = _one_cell(c, ms, q, state(0.0, 0.1, 1e-3, 0.7, 2000.0))[3]
    n45b = _one_cell(c, ms, q, state(0.0, 0.1, 2e-3, 0.7, 2000.0))[3]
    assert (n45b - n5) / (n45 - n5) == pytest.approx(2 ** 0.721, rel=1e-9)


def test_rr2_carries_over_from_the_previous_open_cell(developed):
    """GSLV3M's RR2 (C/Gslv3m.f90:45-60): at a cell without CH4 the O2
    inflow of the CO oxidation is reduced by 1.5 x the RR2 the previous
    open cell (k - 2) left, its NO rate.  With the O2 inflow binding, the
    cell's relaxed CO rate changes by -0.3 * 2 * 1.5 = -0.9 times the change
    of that NO rate; with CH4 present the previous cell does not matter."""
    c = developed
    ms = minor.MinorSpecies(c)
    oracle.initsvm(c, ms)
    t0 = c.ctrl.t0
    i, j, k = c.shape[0] // 2 // 2 * 2, c.shape[1] // 2 // 2 * 2, c.shape[2] // 2 // 2 * 2
    ib = c.arrays["ibcell"]
    assert ib[i - 1, j - 1, k - 1] == 0 and ib[i - 1, j - 1, k - 3] == 0
    g = ms.arrays["gfm"]
    t = c.arrays["t"]
    C, P = (i - 1, j - 1, k - 1), (i - 1, j - 1, k - 3)
    nb = [(i - 3, j - 1, k - 1), (i + 1, j - 1, k - 1), (i - 1, j - 3, k - 1), (i - 1, j + 1, k - 1),
          (i - 1, j - 1, k + 1)]

    def rates(tp, ch4):
        g0, tt = g.copy(), t.copy()
        for n in nb + [P]:  # lean O2, rich CO around the cell
            g[n + (slice(None),)] = 0.0
            g[n + (1,)], g[n + (2,)], g[n + (5,)] = 1e-4, 0.09, 0.9
        g[C + (slice(None),)] = 0.0
        g[C + (0,)], g[C + (1,)], g[C + (2,)], g[C + (5,)] = ch4, 0.2, 0.05, 0.7
        t[C] = 1800.0 / t0
        g[P + (1,)], g[P + (2,)], g[P + (5,)] = 0.2, 0.02, 0.7
        t[P] = tp / t0
        ms.arrays["rrms"][...] = 0.0
        oracle.ms_probe(c, ms, "rates")
        rr = ms.arrays["rrms"]
        out = rr[C[0] // 2, C[1] // 2, C[2] // 2, 1], rr[P[0] // 2, P[1] // 2, P[2] // 2, 3]
        g[...] = g0
        t[...] = tt
        return out

    a_co, a_no = rates(1700.0, 0.0)
    b_co, b_no = rates(1800.0, 0.0)
    assert b_no > 2 * a_no > 0 and a_co > 0 and b_co > 0
    assert (b_co - a_co) == pytest.approx(-0.9 * (b_no - a_no), rel=1e-9)
    a2, _ = rates(1700.0, 0.01)
    b2, _ = rates(1800.0, 0.01)
    assert a2 == b2


def test_uniform_field_is_a_fixed_point(developed):
    """A uniform composition (no CH4, CO or CO2; T = 300 K so that the NO
    rate is below 1e-90) and a uniform soot fraction (oxidation off)
    come out of one GSLV3M unchanged: a_p = sum a_nb, the relaxation and
    the volume terms balance for any flow field."""
    c = cs.furnace_small(nx=24, ny=12, nz=8, seed=3)
    oracle.initsv_tail(c)
    oracle.sphase(c, 20)
    c.arrays["t"][...] = 300.0 / c.ctrl.t0
    ms = minor.MinorSpecies(c, aoxid=0.0, rfm=[0.7] * 7 + [0.3])
    y = np.array([0.0, 0.23, 0.0, 0.0, 0.0, 0.77, 0.0])
    ms.arrays["gfm"][...] = y
    ms.arrays["smf"][...] = 1e-4
    oracle.gslv3m(c, ms)
    cen = _cen(c)
    g = ms.arrays["gfm"][cen]
    assert np.abs(g - y).max() < 1e-14
    assert np.abs(ms.arrays["smf"][cen] - 1e-4).max() < 1e-18
    assert ms.ctrl.bal_ms < 1e-14


def test_species_sum_stays_one_without_clamps():
    """The seven equations share one operator and their sources sum to zero,
    so from mass fractions summing to one everywhere the sum stays one
    (bal_ms at rounding) while no clamp binds: at 1000 K the rates are too
    slow to drive a mass fraction below zero.  With CO above its cap of 0.1
    the balance opens."""
    c = _developed()
    c.arrays["t"][...] = 1000.0 / c.ctrl.t0
    ms = minor.MinorSpecies(c)
    oracle.initsvm(c, ms)
    g = ms.arrays["gfm"]
    cen = (slice(1, None, 2),) * 3
    co = np.where(g[cen + (5,)] > 0.05, 0.01, 0.0)  # some CO, taken from N2
    g[cen + (2,)] = co
    g[cen + (5,)] -= co
    half = (slice(c.shape[0] // 2 | 1, None, 2), slice(1, None, 2), slice(1, None, 2))
    g[half + (5,)] += g[half + (0,)]  # no CH4 in the downstream half: CO oxidation runs there
    g[half + (0,)] = 0.0
    assert np.abs(g[cen].sum(axis=-1) - 1.0).max() < 1e-14
    ref = ms.copy()
    oracle.gslv3m(c, ms)
    rr = ms.arrays["rrms"]
    assert rr[..., 0].max() > 0 and rr[..., 1].max() > 0
    assert np.abs(ms.arrays["gfm"] - ref.arrays["gfm"]).max() > 1e-6
    assert ms.ctrl.bal_ms < 1e-14
    g2 = ref.arrays["gfm"]
    g2[cen + (2,)] = np.where(g2[cen + (0,)] < 0.5, 0.2, 0.0)
    g2[cen + (5,)] -= g2[cen + (2,)]
    oracle.gslv3m(c, ref)
    assert ref.ctrl.bal_ms > 1e-4


def test_co_and_no_caps():
    """GSLV3M caps the solved CO at 0.1 and NO at 0.01 (C/Gslv3m.f90:180-
    183) before relaxing: from 0.3 and 0.05 everywhere, with no transport
    and rates off (300 K), every open cell ends at rfmc * old + rfm * cap."""
    c = _developed(iters=5)
    c.arrays["gama"][...] = 0.0
    c.arrays["ug"][...] = 0.0
    c.arrays["t"][...] = 300.0 / c.ctrl.t0
    ms = minor.MinorSpecies(c, rfm=[0.6] * 7 + [0.2])
    oracle.initsvm(c, ms)
    g = ms.arrays["gfm"]
    g[..., 2] = 0.3
    g[..., 6] = 0.05
    ms2 = ms.copy()
    ms2.arrays["gfm"][..., 2] = 0.05
    ms2.arrays["gfm"][..., 6] = 0.005
    for m in (ms, ms2):
        oracle.gslv3m(c, m)
    cen = _cen(c)
    # without coefficients every row is pinned: the solve returns the old value
    np.testing.assert_allclose(ms.arrays["gfm"][..., 2][cen], 0.4 * 0.3 + 0.6 * 0.1, rtol=1e-15)
    np.testing.assert_allclose(ms.arrays["gfm"][..., 6][cen], 0.4 * 0.05 + 0.6 * 0.01, rtol=1e-15)
    np.testing.assert_allclose(ms2.arrays["gfm"][..., 2][cen], 0.05, rtol=1e-15)
    np.testing.assert_allclose(ms2.arrays["gfm"][..., 6][cen], 0.005, rtol=1e-15)


def test_soot_sources(developed):
    """SOURCEM(11), the new model: formation aform Y_CH4 (exp(-Esf/RT) -
    exp(-Esf/RT0_sf)) r0/(rho0 u0) above T0_sf, oxidation a sink linear in
    O2 above T0_so; after the first minor-species iteration both are
    relaxed against the previous call's (kept whole as the next history)."""
    c = developed
    ms = minor.MinorSpecies(c, aform=0.5, aoxid=30.0, t0_sf=1200.0, t0_so=800.0)
    oracle.initsvm(c, ms)
    cen = _cen(c)
    tk = c.arrays["t"][cen] * c.ctrl.t0
    y1 = ms.arrays["gfm"][..., 0][cen]
    y2 = ms.arrays["gfm"][..., 1][cen]
    ms.ctrl.itr_ms = 1
    oracle.ms_probe(c, ms, "sourcem", 11)
    s = oracle.scratch(c)
    bs, sfp = s["bs"][cen], s["sfp"][cen]
    snorm = c.ctrl.r0 / (c.ctrl.dnst0 * c.ctrl.ug0)
    assert np.all(bs[tk <= 1200.0] == 0.0) and np.all(sfp[tk <= 800.0] == 0.0)
    hot = (tk > 1300.0) & (y1 > 1e-3)
    assert hot.any()
    # linear in Y_CH4 at fixed T: bs / Y_CH4 is the same function of T
    e = np.exp(-ms.ctrl.esf_r / tk[hot]) - np.exp(-ms.ctrl.esf_r / 1200.0)
    np.testing.assert_allclose(bs[hot] / (y1[hot] * e * snorm), 0.5, rtol=1e-13)
    ox = (tk > 900.0) & (y2 > 1e-3)
    assert np.all(sfp[ox] < 0.0)
    f = np.sqrt(tk[ox]) * np.exp(-ms.ctrl.eso_r / tk[ox]) - np.sqrt(800.0) * np.exp(-ms.ctrl.eso_r / 800.0)
    np.testing.assert_allclose(-sfp[ox] / (y2[ox] * f * snorm), 30.0, rtol=1e-13)
    np.testing.assert_array_equal(ms.arrays["bs_soot_prev"], s["bs"])
    np.testing.assert_array_equal(ms.arrays["sfp_soot_prev"], s["sfp"])
    # relaxation after the first iteration
    hist_b = np.full(c.shape, 1e-3, order="F")
    hist_s = np.full(c.shape, -2e-3, order="F")
    ms.arrays["bs_soot_prev"][...] = hist_b
    ms.arrays["sfp_soot_prev"][...] = hist_s
    ms.ctrl.itr_ms = 2
    oracle.ms_probe(c, ms, "sourcem", 11)
    s2 = oracle.scratch(c)
    np.testing.assert_allclose(s2["bs"][cen][hot], 0.3 * bs[hot] + 0.7 * 1e-3, rtol=1e-14)
    np.testing.assert_allclose(s2["sfp"][cen][ox], 0.3 * sfp[ox] + 0.7 * -2e-3, rtol=1e-14)


def test_soot_without_transport_is_formation_over_oxidation():
    """With no flow and no diffusion (every coefficient zero) and soot
    relaxation 1, the soot equation of each cell is sfp_v smf = bs_v with
    both sources per unit volume times the cell volume: smf = formation /
    oxidation rate, whatever the volume (the species rows, without
    coefficients, keep their values)."""
    c = _developed(iters=5)
    c.arrays["gama"][...] = 0.0
    c.arrays["ug"][...] = 0.0
    ms = minor.MinorSpecies(c, rfm=[0.5] * 7 + [1.0], soot_cap=1.0, aform=0.5, aoxid=30.0)
    oracle.initsvm(c, ms)
    ms.ctrl.itr_ms = 1
    probe = ms.copy()
    oracle.ms_probe(c, probe, "sourcem", 11)
    s = oracle.scratch(c)
    cen = _cen(c)
    g0 = ms.arrays["gfm"].copy()
    oracle.gslv3m(c, ms)
    np.testing.assert_array_equal(ms.arrays["gfm"][cen], g0[cen])
    bs, sfp = s["bs"][cen], s["sfp"][cen]
    ok = sfp < -1e-10  # rows with a_p below 1e-20 keep their value
    assert ok.sum() > 100
    want = bs[ok] / -sfp[ok]
    want = np.where(want > 1.0, 1.0, np.where(want > 1e-20, want, 0.0))  # clipped at soot_cap
    np.testing.assert_allclose(ms.arrays["smf"][cen][ok], want, rtol=1e-13, atol=1e-300)


def test_intpm_is_intp_on_the_same_values(developed):
    """intpm repeats INTP's loop nest and weights (with INTP0's in-place
    inlet weight): with the major species in GFM's slots, the face values of
    the two agree to rounding (INTP is pinned by the linear-field tests)."""
    c = _developed(iters=5)
    ms = minor.MinorSpecies(c)
    oracle.initsvm(c, ms)
    oracle.probe(c, "intp")
    g = ms.arrays["gfm"]
    GF = c.arrays["gf"]
    g[...] = 0.0
    cen = (slice(1, None, 2),) * 3
    pairs = ((0, cs.IYF), (1, cs.IYO2), (3, cs.IYH2O), (4, cs.IYCO2), (5, cs.IYN2))
    for s, l in pairs:
        g[cen + (s,)] = GF[cen + (l - 1,)]
    oracle.ms_probe(c, ms, "intpm")
    ib = c.arrays["ibcell"]
    face = np.ones(ib.shape, dtype=bool)
    face[cen] = False
    face &= (ib != 1) & (ib != 3)
    face[0, :, :] = face[:, 0, :] = face[:, :, 0] = False
    # INTP0 also renormalises through CONC (H2O = 1 - the others, or all
    # divided by their sum above one), a change at the rounding level
    for s, l in pairs:
        np.testing.assert_allclose(g[..., s][face], GF[..., l - 1][face], rtol=0, atol=1e-15)


def test_extrm_exit_extrapolates_linear_fields_and_walls_copy(developed):
    c = _developed(nx=16, ny=8, nz=8, iters=2)
    ms = minor.MinorSpecies(c)
    oracle.initsvm(c, ms)
    A = c.arrays
    x = A["x"]
    g = ms.arrays["gfm"]
    ib = A["ibcell"]
    lin = 0.01 + 0.02 * x / x.max()
    g[...] = lin[:, None, None, None]
    ms.arrays["smf"][...] = np.arange(c.shape[0])[:, None, None] * 1e-5
    before = g.copy()
    oracle.ms_probe(c, ms, "extrm")
    cen = np.zeros(ib.shape, dtype=bool)
    cen[1::2, 1::2, 1::2] = True
    ex = cen & (ib % 10 == 3)
    assert ex.any()
    I = np.nonzero(ex)[0]
    # the exits of furnace_small are x-normal at the x = max end
    np.testing.assert_allclose(g[ex][:, 0], lin[I], rtol=1e-13)
    np.testing.assert_array_equal(ms.arrays["smf"][ex], (I - 2) * 1e-5)
    wall = cen & (ib % 10 == 1)
    assert np.all(np.isin(g[wall][:, 0], lin))
    oth = cen & ~ex & ~wall
    np.testing.assert_array_equal(g[oth], before[oth])


def test_msflow_convergence_test_and_maxms(developed):
    """MSFLOW's loop (C/Msflow.f90:48-162): itr_ms += 1, GSLV3M, the test
    (itr_ms > 10 and every resid_ms <= bmcon) and, unless it ends the
    loop, extrm and intpm; maxms bounds the iterations.  (A GSLV3M first
    puts the species coefficients into AS, which the rates of the first
    GSLV3M of each sequence read.)"""
    c = developed
    warm = minor.MinorSpecies(c)
    oracle.initsvm(c, warm)
    oracle.gslv3m(c, warm)
    ms = minor.MinorSpecies(c, maxms=50, bmcon=1e30)
    oracle.initsvm(c, ms)
    ref = ms.copy()
    oracle.msflow(c, ms)
    assert ms.ctrl.iterations == 11 and ms.ctrl.converged == 1 and ms.ctrl.itr_ms == 11
    # the same by hand: GSLV3M; 10 x (extrm, intpm, GSLV3M), nothing after the last
    ref.ctrl.igs0 = 0
    for n in range(11):
        ref.ctrl.itr_ms += 1
        oracle.gslv3m(c, ref)
        if n < 10:
            oracle.ms_probe(c, ref, "extrm")
            oracle.ms_probe(c, ref, "intpm")
    for k in ms.arrays:
        np.testing.assert_array_equal(ms.arrays[k], ref.arrays[k])
    ms2 = minor.MinorSpecies(c, maxms=4)
    oracle.initsvm(c, ms2)
    oracle.msflow(c, ms2)
    assert ms2.ctrl.iterations == 4 and ms2.ctrl.converged == 0 and ms2.ctrl.itr_ms == 4
