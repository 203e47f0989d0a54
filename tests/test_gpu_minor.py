"""GPU parity of the minor-species solve (INITSVM, GSLV3M, MSFLOW; C/Initsv.f90:
329-380, C/Gslv3m.f90, C/Msflow.f90) against the oracle.

Both sides start from one gas state: the oracle's, after a few SIMPLER
iterations, uploaded to the device.  The rates of the first GSLV3M after a
GFLOW read the coefficients AS the last GFLOW left; the two sides' AS agree
only to rounding after their own GFLOWs, so each test first runs one
GSLV3M on both (which puts the species coefficients, assembled from the
shared state, into AS) and then starts the compared sequence from an
uploaded minor-species state.  Compared: GFM, smf, the soot-source history
and RRMS over the whole lattice, and the control block's outputs.

  * reference plane order, FP64: GSLV3M x 3 and MSFLOW (15 iterations, with
    extrm and intpm between) to 1e-10 relative;
  * the restart branch of INITSVM (inlet override, extrm, intpm);
  * the stretched furnace with burners on the y walls and a melt floor;
  * zebra plane order and FP32 sweeps: 40 MSFLOW iterations against the
    oracle's in the reference order (the minor-species iteration amplifies
    rounding differences: compared trajectories stay short).
"""
import numpy as np
import pytest

import oracle
from gfm_amd import case as cs
from gfm_amd import minor
from gfm_amd.solver import CombustionSpace

pytestmark = pytest.mark.gpu

OUTS = ("bal_ms", "soot_tot", "smf_max", "smf_max_noclip", "rr_f", "rr_c")


def _gas(base, iters=20):
    o = base.clone()
    oracle.initsv_tail(o)
    oracle.sphase(o, iters)
    return o


def _ms(case, **kw):
    kw.setdefault("aform", 0.00353)
    kw.setdefault("aoxid", 30.0)
    return minor.MinorSpecies(case, **kw)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(a).max(), np.abs(b).max(), 1e-300))


def _compare(mo, mg, tol, what="", outs=True):
    errs = {}
    for k in ("smf", "bs_soot_prev", "sfp_soot_prev"):
        errs[k] = _rel(mo.arrays[k], mg.arrays[k])
    for l in range(7):
        errs[f"gfm{l + 1}"] = _rel(mo.arrays["gfm"][..., l], mg.arrays["gfm"][..., l])
    for r in (0, 1, 3):
        errs[f"rrms{r + 1}"] = _rel(mo.arrays["rrms"][..., r], mg.arrays["rrms"][..., r])
    print(what, {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v < tol, (what, k, v)
    if not outs:
        return
    for k in OUTS:
        a, b = getattr(mo.ctrl, k), getattr(mg.ctrl, k)
        print(f"  {k:15s} oracle {a:.16e} gpu {b:.16e}", flush=True)
        floor = 1e-14 if k == "bal_ms" else 1e-300  # |1 - sum of 7 mass fractions|: rounding where no clamp binds
        assert abs(a - b) <= tol * max(abs(a), abs(b), 1e-300) + floor, (what, k, a, b)
    assert mo.ctrl.smf_clipped_cnt == mg.ctrl.smf_clipped_cnt
    for l in range(8):
        a, b = mo.ctrl.resid_ms[l], mg.ctrl.resid_ms[l]
        assert abs(a - b) <= max(1e-6 * max(a, b), 1e-15), (what, "resid_ms", l, a, b)


def _start(o, s, **kw):
    """Warm-up GSLV3M on both sides (species AS into AS), then a fresh
    INITSVM from zero rates on both: returns (oracle state, device state)."""
    w = _ms(o, **kw)
    oracle.initsvm(o, w)
    oracle.gslv3m(o, w)
    wg = _ms(o, **kw)
    s.ms_upload(wg)
    s.initsvm(wg)
    s.gslv3m(wg)
    mo = _ms(o, **kw)
    oracle.initsvm(o, mo)
    mg = _ms(o, **kw)
    s.ms_upload(mg)
    s.initsvm(mg)
    return mo, mg


@pytest.mark.parametrize("geom", ["uniform", "stretched"])
def test_gslv3m_and_msflow_match_oracle(geom):
    base = cs.furnace_small(nx=24, ny=12, nz=8, seed=3) if geom == "uniform" else \
        cs.furnace_large(nx=32, ny=24, nz=12, seed=3)
    base.ctrl.sweep_order = 0
    o = _gas(base)
    g = o.clone()
    with CombustionSpace(g, initsv=False) as s:
        mo, mg = _start(o, s)
        s.ms_download(mg)
        _compare(mo, mg, 1e-13, f"{geom} INITSVM", outs=False)
        for n in range(3):
            mo.ctrl.itr_ms += 1
            mg.ctrl.itr_ms += 1
            oracle.gslv3m(o, mo)
            s.gslv3m(mg)
            s.ms_download(mg)
            _compare(mo, mg, 1e-10, f"{geom} GSLV3M {n + 1}")
        mo.ctrl.maxms = mg.ctrl.maxms = 15
        mo.ctrl.bmcon = mg.ctrl.bmcon = 1e-30  # no early exit
        oracle.msflow(o, mo)
        s.msflow(mg)
        s.ms_download(mg)
        assert mo.ctrl.iterations == mg.ctrl.iterations == 15 and mg.ctrl.itr_ms == 18
        _compare(mo, mg, 1e-9, f"{geom} MSFLOW")
        for l in range(7):
            assert mg.ctrl.gfm_0[l] == pytest.approx(mo.ctrl.gfm_0[l], rel=1e-9, abs=1e-18)
        assert mg.ctrl.gfmex == pytest.approx(mo.ctrl.gfmex, rel=1e-9)
        assert mg.ctrl.svf_r == pytest.approx(mo.ctrl.svf_r, rel=1e-9)
        # restart: the record written back, INITSVM's restart branch on both
        ro, rg = mo.copy(), mg.copy()
        ro.arrays["gfm"][1::2, 1::2, 1::2, :] *= 0.9  # a record that differs from the running state
        rg.arrays["gfm"][...] = ro.arrays["gfm"]
        oracle.initsvm(o, ro, restart=True)
        s.ms_upload(rg)
        s.initsvm(rg, restart=True)
        s.ms_download(rg)
        _compare(ro, rg, 1e-9, f"{geom} restart INITSVM", outs=False)


@pytest.mark.parametrize("order,precision,tol", [(1, 0, 2e-6), (1, 1, 1e-4)])
def test_msflow_in_zebra_order_and_fp32(order, precision, tol):
    """40 MSFLOW iterations on the uniform chamber in zebra plane order
    (FP64) and with FP32 sweeps against the oracle's in the reference order.
    Each species solve of this case ends at calc_residual 1e-16 (the soot
    row of an isolated corner cell without coefficients keeps the soot
    residual at 1/NCELLS), so the plane order changes a solve at the 1e-16
    level.  The iteration itself amplifies differences: in the oracle a
    1e-13 change of CH4 grows to 1e-10 by iteration 100 and to O(1) in NO by
    iteration 160 (hot cells, NO and CO), so trajectories are compared over
    40 iterations."""
    base = cs.furnace_small(nx=24, ny=12, nz=8, seed=3)
    base.ctrl.sweep_order = 0
    o = _gas(base)
    g = o.clone()
    g.ctrl.sweep_order = order
    g.ctrl.precision = precision
    with CombustionSpace(g, initsv=False) as s:
        mo, mg = _start(o, s)
        mo.ctrl.maxms = mg.ctrl.maxms = 40
        mo.ctrl.bmcon = mg.ctrl.bmcon = 1e-30
        oracle.msflow(o, mo)
        s.msflow(mg)
        s.ms_download(mg)
    assert mg.ctrl.iterations == 40
    cen = np.zeros(o.shape, dtype=bool)
    cen[3:-2:2, 3:-2:2, 3:-2:2] = o.arrays["ibcell"][3:-2:2, 3:-2:2, 3:-2:2] <= 0
    for l in range(7):
        a, b = mo.arrays["gfm"][..., l][cen], mg.arrays["gfm"][..., l][cen]
        e = float(np.abs(a - b).max() / max(np.abs(a).max(), 1e-300))
        print(f"  gfm{l + 1} rel {e:.2e}", flush=True)
        assert e < tol, (l, e)
    a, b = mo.arrays["smf"][cen], mg.arrays["smf"][cen]
    e = float(np.abs(a - b).max() / max(np.abs(a).max(), 1e-300))
    print(f"  smf rel {e:.2e}", flush=True)
    assert e < tol
