"""GPU parity of the wall radiosity exchange (C/Qwall.f90: the patch list
and view factors of prepare_for_wall_radiation_calculation, wall_q_amb,
wallq with calc_q_conv_wall; the first-pass patch temperatures of
RAD_EMIS; QRN_PRN's melt sums) and of the state it hands the gas
iterations (T and h of the wall cells, the patch q_conv as the enthalpy
sink of C/Source.f90:473-509), against the oracle.

  * the reference runs' saved states (case0041, case1206): after the
    device's volume pass, the device's exchange reproduces the heat to the
    melt and the mean wall temperature their last pass logged (to 1e-12),
    the oracle's patch list exactly, its view factors to 1e-15, and every
    patch's radiosity, temperature, q_conv, q_amb and qc and the melt flux
    to 1e-12 -- the sweeps follow the reference's patch order and add the
    other patches' radiation in another order, so the sweep counts agree
    and each sweep's mean relative change to 1e-6 of itself;
  * a developed stretched furnace with a melt floor and an exit port:
    volume pass, exchange, then two SIMPLER iterations on both sides with
    QA - QE and the patch sink (fields to 1e-9, as the other
    developed-state comparisons).
"""
import os
import time

import numpy as np
import pytest

import oracle
from gfm_amd import case as cs
from gfm_amd import minor, radiation
from gfm_amd import wall as wl
from gfm_amd.abi import RADSRC_FIELD, WALL_I16
from gfm_amd.solver import CombustionSpace
from helpers import field_errors, wall_reference_state

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
PATCH = ("radiosity", "b_t", "q_conv", "q_amb", "qc", "qeb")


def _cmp(a, b, tol, what):
    scale = max(np.abs(a).max(), np.abs(b).max(), 1e-300)
    e = float(np.abs(a - b).max() / scale)
    print(f"  {what:10s} max |diff|/max {e:.2e}", flush=True)
    assert e <= tol, (what, e)


@pytest.mark.parametrize("cn", ["0041", "1206"])
def test_exchange_matches_the_log_and_the_oracle(cn):
    c, ms, rad, wc, iwall, fx = wall_reference_state(cn, GOLD)
    co, ro, wo = c.clone(), rad.copy(), wl.Wall(c, wc.copy())
    print(f"case{cn}: oracle pass and exchange", flush=True)
    t0 = time.perf_counter()
    oracle.rad_emis(co, ms, ro)
    t1 = time.perf_counter()
    oracle.wall_setup(co, wo, iwall)
    oracle.wallq(co, wo, ro, iwall)
    t2 = time.perf_counter()
    print(f"case{cn}: device (oracle: volume {t1 - t0:.2f} s, exchange {t2 - t1:.2f} s)", flush=True)
    cg = c.clone()
    with CombustionSpace(cg, initsv=False) as s:
        s.ms_upload(ms)
        rg = rad.copy()
        s.rad_emis(rg, download=False)
        wg = wl.Wall(cg, wc.copy())
        t3 = time.perf_counter()
        s.wall_setup(wg, iwall)
        t4 = time.perf_counter()
        s.wallq(wg)
        t5 = time.perf_counter()
        print(f"case{cn}: device setup {t4 - t3:.3f} s, exchange {t5 - t4:.3f} s ({wg.ctrl.sweeps} sweeps, "
              f"{wg.ctrl.outer} outer iterations; oracle {wo.ctrl.sweeps}, {wo.ctrl.outer})", flush=True)
        s.download()
    k, ko = wg.ctrl, wo.ctrl
    print(f"  q_melt log {fx['q_melt']:.14e} device {k.q_melt:.14e} oracle {ko.q_melt:.14e}", flush=True)
    print(f"  mean wall T log {fx['avg_wall_T']:.14e} device {k.avg_wall_T:.14e} oracle {ko.avg_wall_T:.14e}",
          flush=True)
    assert k.q_melt == pytest.approx(fx["q_melt"], rel=1e-12)
    assert k.avg_wall_T == pytest.approx(fx["avg_wall_T"], rel=1e-12)
    assert (k.b_cnt, k.iwall_count, k.imelt_count, k.inlet_count, k.iexit_count) == \
        (ko.b_cnt, ko.iwall_count, ko.imelt_count, ko.inlet_count, ko.iexit_count)
    for key in ("wall_area_total", "melt_area_total", "inlet_area_total", "exit_area_total"):
        assert getattr(k, key) == getattr(ko, key), key
    for key in WALL_I16 + ("patch_area", "b_face_area", "emis"):
        assert np.array_equal(wg.arrays[key], wo.arrays[key]), key
    _cmp(wg.arrays["vf_sum"], wo.arrays["vf_sum"], 1e-15, "vf_sum")
    _cmp(wg.arrays["vf"], wo.arrays["vf"], 1e-15, "vf")
    for key in PATCH:
        _cmp(wg.arrays[key], wo.arrays[key], 1e-12, key)
    _cmp(wg.arrays["qrs"], wo.arrays["qrs"], 1e-12, "qrs")
    _cmp(wg.arrays["wall_rad_t"], wo.arrays["wall_rad_t"], 1e-12, "wall_rad_t")
    assert k.sweeps == ko.sweeps and k.outer == ko.outer and k.iterw == ko.iterw
    np.testing.assert_allclose(wg.log(), wo.log(), rtol=1e-6, atol=1e-16)
    for key in ("q_incident_total", "qconv_total", "q_wall_loss", "q_rad_inlet", "q_rad_outlet", "qm_flame",
                "qm_bnds", "qm_conv", "qa_s0", "aform"):
        assert getattr(k, key) == pytest.approx(getattr(ko, key), rel=1e-11, abs=1e-9), key
    # the state of the next gas iterations: T at wall faces and cells, h at wall cells
    m = (wo.arrays["wall_rad_t"] > 0)
    _cmp(cg.arrays["t"][m], co.arrays["t"][m], 1e-12, "wall T")
    _cmp(cg.arrays["gf"][..., 2][m], co.arrays["gf"][..., 2][m], 1e-12, "wall h")


def _exhaust_of(case, wc):
    """The exit port of a furnace_large case as its one exhaust (fixed
    temperature), in the sbc file's terms (C/Sbc.f90:272-284)."""
    ib = case.arrays["ibcell"]
    mp = ib.shape[0]
    jj, kk = np.nonzero(ib[mp - 1] == 3)
    jj, kk = jj + 1, kk + 1
    wc.nex = 1
    for q, v in enumerate((mp, int(jj.min()), int(kk.min()), mp, int(jj.max()), int(kk.max()))):
        wc.exh_box[0][q] = v
    wc.exh_orient[0], wc.exh_direct[0], wc.exh_type[0] = 1, 1, 1
    wc.exh_frac[0], wc.exh_fixed[0] = 0.9, 1500.0


def test_gas_iterations_after_the_exchange():
    base = cs.furnace_large(nx=32, ny=12, nz=8, seed=4, walls="adiabatic", gray_kappa=0.0, exit_port=(1.0, 2.0, 0.4, 1.2))
    base.ctrl.sweep_order = 0
    o = base.clone()
    oracle.initsv_tail(o)
    oracle.sphase(o, 20)
    ms = minor.MinorSpecies(o, aform=0.00353, aoxid=30.0)
    oracle.initsvm(o, ms)
    rad = radiation.Radiation(o, wl=[x * 1e-6 for x in (1, 2, 3, 4, 5, 6, 7, 8.5, 12, 15)], eps_m=0.9,
                              have_crown=1, height_to_ceiling=1.2)
    wc = wl.default_ctrl(eps_m=0.9, have_crown=1, height_to_ceiling=1.2)
    _exhaust_of(o, wc)
    g = o.clone()
    ro, wo = rad.copy(), wl.Wall(o, wc.copy())
    oracle.rad_emis(o, ms, ro)
    oracle.wall_setup(o, wo, None)
    oracle.wallq(o, wo, ro, None)
    assert wo.ctrl.imelt_count > 0 and wo.ctrl.iexit_count > 0
    with CombustionSpace(g, initsv=False) as s:
        s.ms_upload(ms)
        rg = rad.copy()
        s.rad_emis(rg)
        wg = wl.Wall(g, wc.copy())
        s.wall_setup(wg, None)
        s.wallq(wg)
        _cmp(rg.arrays["qa"], ro.arrays["qa"], 1e-12, "qa")
        for key in PATCH:
            _cmp(wg.arrays[key], wo.arrays[key], 1e-12, key)
        # both sides: QA - QE of the oracle's pass, the patch sink, walls no longer adiabatic
        s.rad_upload(ro)
        oracle.lib().oracle_rad_field(oracle._dp(ro.arrays["qa"]), oracle._dp(ro.arrays["qe"]))
        oracle.wall_field(wo.qconv_face())
        try:
            for case in (o, g):
                case.ctrl.gray_radiation = RADSRC_FIELD
                case.ctrl.i1st_rad_done = 1
            oracle.sphase(o, 2)
            s.sphase(2)
            s.download()
        finally:
            oracle.lib().oracle_rad_field(None, None)
            oracle.wall_field(None)
    errs = field_errors(o, g)
    print("  fields after 2 iterations with the patch sink:", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v < 1e-9, (k, v)
    assert abs(o.ctrl.q_wall - g.ctrl.q_wall) <= 1e-9 * max(abs(o.ctrl.q_wall), 1.0)
