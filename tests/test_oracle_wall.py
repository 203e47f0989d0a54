"""Pins of the oracle's wall radiosity exchange (C/Qwall.f90:
prepare_for_wall_radiation_calculation, wall_q_amb, wallq with
calc_q_conv_wall and the preset view factors; the first-pass patch
temperatures of RAD_EMIS; QRN_PRN's melt sums) and of the radiation
restart record rr####c.d.

  the whole exchange    two reference runs (case0041 and case1206): after
                        the volume pass on the state of their saved
                        records, the exchange reproduces what their last
                        radiation pass logged (tests/golden/
                        make_wall_fixtures.py): the heat to the melt and
                        the mean wall temperature to 1e-12, the wall
                        temperature of every wall cell to the 7 digits
                        printed, the melt-surface heat flux of every melt
                        cell to the 4 digits printed, the patch and
                        wall-patch counts, and (case1206, soot
                        calibration) the adjusted soot formation constant
                        to 1e-12.
  the converged state   every patch satisfies the equation it was iterated
                        on, evaluated here in numpy from the arrays: a
                        wall's radiosity is qc + q_conv - q_amb + the
                        radiosity it sees, an inlet's, exit's or melt
                        patch's is its emission plus the reflected part;
                        wall temperatures invert the radiosity.
  geometry              each row of view factors sums to one; the area of
                        the patches on a boundary cell equals the sum of
                        their patch areas (grouped in numpy); patch areas
                        are the face areas of their normal.
  rr record             written and read back byte for byte.
"""
import os

import numpy as np
import pytest

import oracle
from gfm_amd import wall as wl
from helpers import wall_reference_state

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
SIG = 5.668e-8
_CACHE = {}


def run(cn):
    if cn not in _CACHE:
        c, ms, rad, wc, iwall, fx = wall_reference_state(cn, GOLD)
        oracle.rad_emis(c, ms, rad)
        w = wl.Wall(c, wc)
        oracle.wall_setup(c, w, iwall)
        oracle.wallq(c, w, rad, iwall)
        _CACHE[cn] = (c, rad, w, fx)
    return _CACHE[cn]


@pytest.mark.parametrize("cn", ["0041", "1206"])
def test_exchange_reproduces_the_reference_logs(cn):
    c, rad, w, fx = run(cn)
    k = w.ctrl
    print(f"  q_melt {k.q_melt:.14e} (log {fx['q_melt']:.14e}); mean wall T {k.avg_wall_T:.14e} "
          f"(log {fx['avg_wall_T']:.14e}); {k.sweeps} sweeps, {k.outer} outer iterations")
    assert k.b_cnt == fx["b_cnt"] and k.iwall_count == fx["iwall_count"]
    assert k.q_melt == pytest.approx(fx["q_melt"], rel=1e-12)
    assert k.avg_wall_T == pytest.approx(fx["avg_wall_T"], rel=1e-12)
    wt = w.arrays["wall_rad_t"]
    ref = np.zeros_like(wt)
    for i, j, kk, v in fx["wall_rad_T"]:
        ref[i - 1, j - 1, kk - 1] = v
    assert np.array_equal(wt > 0, ref > 0)
    m = ref > 0
    assert np.max(np.abs(wt[m] - ref[m]) / ref[m]) <= 5.01e-7
    qrs = w.arrays["qrs"]
    assert np.count_nonzero(qrs) == len(fx["qrs"])
    for i, j, v in fx["qrs"]:
        assert abs(qrs[i - 1, j - 1] - v) <= 5.01e-4 * abs(v), (i, j)
    if "aform" in fx:
        assert k.aform == pytest.approx(fx["aform"], rel=1e-12)
    assert k.rel_rad_chg_last < 1e-13


@pytest.mark.parametrize("cn", ["0041", "1206"])
def test_converged_patches_satisfy_their_equations(cn):
    c, rad, w, fx = run(cn)
    a = w.arrays
    J, vf, e, t = a["radiosity"], a["vf"], a["emis"], a["b_typ"]
    H = vf @ J
    wall = t == 1
    rhs = np.where(wall, np.maximum(a["qc"] + a["q_conv"] - a["q_amb"] + H, e * SIG * 200.0 ** 4),
                   a["qeb"] + (1.0 - e) * (np.where((t == 1) | (e < 1.0), H, 0.0) + a["qc"]))
    assert np.max(np.abs(rhs - J) / J) < 1e-12
    sig_t4 = J[wall] - (a["q_amb"][wall] - a["q_conv"][wall]) * (1.0 - e[wall]) / e[wall]
    np.testing.assert_allclose(SIG * a["b_t"][wall] ** 4, np.maximum(sig_t4, SIG * 200.0 ** 4), rtol=1e-12)
    melt = t == 4
    np.testing.assert_allclose(a["qeb"][melt], e[melt] * SIG * a["b_t"][melt] ** 4, rtol=1e-15)
    qm = np.sum((a["qc"] + a["q_conv"] + H - J)[melt] * a["patch_area"][melt])
    assert qm == pytest.approx(w.ctrl.q_melt, rel=1e-12)


@pytest.mark.parametrize("cn", ["0041", "1206"])
def test_patch_geometry(cn):
    c, rad, w, fx = run(cn)
    a = w.arrays
    vf = a["vf"]
    rs = vf.sum(axis=1)
    assert np.all(a["vf_sum"] > 0.0)
    np.testing.assert_allclose(rs, 1.0, atol=1e-13)
    assert np.all(np.diag(vf) == 0.0)
    key = (a["bc_i"].astype(np.int64) * 1000 + a["bc_j"]) * 1000 + a["bc_k"]
    _, inv = np.unique(key, return_inverse=True)
    tot = np.bincount(inv, weights=a["patch_area"])
    np.testing.assert_allclose(a["b_face_area"], tot[inv], rtol=1e-15)
    x, dx, dr, dz = (c.arrays[k] for k in ("x", "dx", "dr", "dz"))
    i, j, k, o = (a[q].astype(int) - 1 for q in ("b_i", "b_j", "b_k", "b_orient"))
    want = np.where(o == 0, dr[j] * dz[k], np.where(o == 1, dx[i] * dz[k], dx[i] * dr[j]))
    assert np.array_equal(a["patch_area"], want)
    # the boundary cell is the patch's neighbour away from the open side; an exit patch sits where its
    # tunnel meets the interior, on a cell the grid may leave open
    ib = c.arrays["ibcell"]
    bc = (a["bc_i"] - 1, a["bc_j"] - 1, a["bc_k"] - 1)
    assert np.all((ib[bc] >= 1) | (a["b_typ"] == 3))
    step = np.where(a["b_direct"] == 0, -1, 1)
    for q, name in enumerate(("bc_i", "bc_j", "bc_k")):
        d = a[name] - a[("b_i", "b_j", "b_k")[q]]
        assert np.array_equal(d, np.where(a["b_orient"] == q + 1, step, 0)), name


def test_rr_record_round_trip(tmp_path):
    c, rad, w, fx = run("0041")
    ql = np.zeros_like(rad.arrays["qa"])
    sums = {"qeg0": rad.ctrl.qeg0, "qeg_g": rad.ctrl.qeg_g, "qeg_w": rad.ctrl.qeg_w, "qeg_s": rad.ctrl.qeg_s,
            "qeg_io": rad.ctrl.qeg_io, "q_exhaust": 1234.5, "aoxid": 30.0}
    raw = wl.rr_bytes(rad.arrays["qe"], rad.arrays["qa"], ql, w, sums)
    n = w.ctrl.b_cnt
    assert len(raw) > 8 * n * n  # the view factors are in the record
    p = tmp_path / "rr0041c.d"
    wl.write_rr(p, rad.arrays["qe"], rad.arrays["qa"], ql, w, sums)
    qe, qa, ql2, w2, s2 = wl.read_rr(str(p), c)
    assert np.array_equal(qe, rad.arrays["qe"]) and np.array_equal(qa, rad.arrays["qa"])
    for key in wl.WALL_I16 + wl.WALL_F64 + ("vf", "vf_sum", "wall_rad_t"):
        assert np.array_equal(w2.arrays[key], w.arrays[key]), key
    assert w2.ctrl.avg_wall_T == w.ctrl.avg_wall_T and w2.ctrl.iterw == w.ctrl.iterw
    assert s2["q_exhaust"] == 1234.5 and s2["aoxid"] == 30.0
    assert wl.rr_bytes(qe, qa, ql2, w2, s2) == raw
