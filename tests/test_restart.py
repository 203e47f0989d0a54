"""Restart records rg####c.d (gfm_amd/restart.py) against the reference's
own restart files and the read/write rules of SAV3F (C/Sav3f.f90:21-88)
and read_restart_data (C/Initsv.f90:225-316)."""
import glob
import gzip
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from gfm_amd import case as cs
from gfm_amd import restart as rs
from helpers import open_centres

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
REF = "/root/reference/GFM_Source_Package_2015-08/combustion"


def _rg0041():
    meta = json.load(open(os.path.join(GOLD, "restart_fixtures.json")))["rg0041c.d"]
    raw = gzip.open(os.path.join(GOLD, "rg0041c.d.gz")).read()
    return meta, raw


def test_reference_restart_file_round_trips():
    """The TC21 case's restart record (case0041, lattice 40x48x36): parsed
    and written back it is the same file, byte for byte."""
    meta, raw = _rg0041()
    assert hashlib.sha256(raw).hexdigest() == meta["sha256"]
    rec = rs.read_records(raw, meta["mp"], meta["np"], meta["lp"], meta["lend"])
    assert rec.cell6.shape == (20 * 24 * 18, 6) and rec.gf.shape == (20 * 24 * 18, 7)
    assert rs.records_bytes(rec) == raw
    # the counters of the run that wrote it (its convg log ends at 7540)
    assert (rec.itr_gas, rec.itr_rad, rec.itr_gend, rec.interval_rad) == (7540, 64, 7600, 100)
    assert rec.i1st_rad_done == 1 and rec.i1st_rad_qe_done == 1
    assert 0.8 <= rec.factor <= 1.5  # FACTOR is clamped to [0.8, 1.5], C/Gslv3v.f90:210-234
    assert rec.t_surf_prev.shape == (20, 24)


def _gd_dims(path):
    v = open(path).read().splitlines()[2].split()
    return int(v[1]), int(v[2]), int(v[3])


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference data files not mounted")
def test_every_reference_restart_file_round_trips():
    """Every rg####c.d the reference ships, with the lattice of its gd file."""
    n = 0
    for rg in sorted(glob.glob(os.path.join(REF, "case*", "rg*c.d"))):
        d = os.path.dirname(rg)
        gd = glob.glob(os.path.join(d, "gd*c.dat"))
        if not gd:
            continue
        raw = open(rg, "rb").read()
        rec = rs.read_records(raw, *_gd_dims(gd[0]))
        assert rs.records_bytes(rec) == raw, rg
        n += 1
    assert n >= 10


def test_corrupt_records_are_rejected():
    meta, raw = _rg0041()
    dims = (meta["mp"], meta["np"], meta["lp"])
    bad = bytearray(raw)
    bad[120 * 100] ^= 1  # the leading marker of cell 100's first record
    with pytest.raises(ValueError):
        rs.read_records(bytes(bad), *dims)
    with pytest.raises(ValueError):
        rs.read_records(raw[:-3], *dims)
    with pytest.raises(ValueError):
        rs.read_records(raw + b"\0" * 8, *dims)
    with pytest.raises(ValueError):
        rs.read_records(raw, dims[0] + 2, dims[1], dims[2])


def _developed(n=4):
    c = cs.furnace_small(nx=16, ny=12, nz=10, seed=3)
    c.ctrl.sweep_order = 1
    oracle.initsv_tail(c)
    oracle.sphase(c, n)
    return c


def test_save_then_load_restores_the_saved_fields(tmp_path):
    """SAV3F then read_restart_data on a fresh build of the same case (a
    real restart, itr_gas = 0): P everywhere; T, THETA, GF(1:7) and the
    +faces of U, V, W at every non-inlet centre; k and eps only at inlet
    cells; rho and Gamma re-derived (DNST_CALC, THER) equal the saved ones
    at open cells; the counters come from the file and maxgi grows by
    itr_gas."""
    a = _developed()
    a.ctrl.factor = 1.0123
    path = tmp_path / "rg0001c.d"
    rs.save_rg(a, path)
    b = cs.furnace_small(nx=16, ny=12, nz=10, seed=3)
    maxgi0 = b.ctrl.maxgi
    b.ctrl.itr_gas = 0
    rs.load_rg(b, path)
    ib = a.arrays["ibcell"]
    cen = np.zeros(ib.shape, bool)
    cen[1::2, 1::2, 1::2] = True
    inlet = cen & (ib == 2)
    rest = cen & ~inlet
    np.testing.assert_array_equal(b.arrays["p"][cen], a.arrays["p"][cen])
    for k in ("t", "theta"):
        np.testing.assert_array_equal(b.arrays[k][rest], a.arrays[k][rest])
    for l in range(7):
        want = a.arrays["gf"][..., l][rest]
        if l == cs.IEPS - 1:  # THER keeps eps >= 1e-12 (wall cells hold 0)
            want = np.maximum(want, 1e-12)
        np.testing.assert_array_equal(b.arrays["gf"][..., l][rest], want)
    for l in (cs.IK, cs.IEPS):
        np.testing.assert_array_equal(b.arrays["gf"][..., l - 1][inlet], a.arrays["gf"][..., l - 1][inlet])
    for d in range(3):
        face = np.zeros(ib.shape, bool)
        sl = [slice(1, None, 2)] * 3
        sl[d] = slice(2, None, 2)
        face[tuple(sl)] = True
        face &= (ib != 1) & (ib != 2)
        np.testing.assert_array_equal(b.arrays["ug"][..., d][face], a.arrays["ug"][..., d][face])
    m = open_centres(a)
    for k in ("dnst", "gama", "tmu"):
        np.testing.assert_allclose(b.arrays[k][m], a.arrays[k][m], rtol=1e-13)
    np.testing.assert_array_equal(b.arrays["bs_hprev"], a.arrays["bs_hprev"])
    assert b.ctrl.itr_gas == a.ctrl.itr_gas == 4
    assert b.ctrl.factor == 1.0123
    assert b.ctrl.maxgi == maxgi0 + 4
    assert b.ctrl.itr_gend == a.ctrl.itr_gend > 4  # ahead of itr_gas, same interval: kept
    rec = rs.read_records(path, *a.shape)
    rec.itr_gend = 2  # behind itr_gas: a new radiation interval starts at itr_gas
    rs.write_records(path, rec)
    c = cs.furnace_small(nx=16, ny=12, nz=10, seed=3)
    c.ctrl.itr_gas = 0
    rs.load_rg(c, path)
    assert c.ctrl.itr_gend == 4 + rs.INTERVAL_RAD


def test_loop_forced_restart_keeps_the_running_counters(tmp_path):
    """With itr_gas > 0 (the reference's restart at every global
    iteration) the counters are not taken from the file and inlet cells
    are read like every other cell."""
    a = _developed(3)
    path = tmp_path / "rg0001c.d"
    rs.save_rg(a, path)
    rec = rs.read_records(path, *a.shape)
    rec.itr_gas = 999
    rs.write_records(path, rec)
    b = a.clone()
    b.ctrl.itr_gas = 7
    b.ctrl.itr_gend = 50
    rs.load_rg(b, path)
    assert b.ctrl.itr_gas == 7 and b.ctrl.itr_gend == 50
    inlet = np.zeros(a.arrays["ibcell"].shape, bool)
    inlet[1::2, 1::2, 1::2] = a.arrays["ibcell"][1::2, 1::2, 1::2] == 2
    assert inlet.any()
    np.testing.assert_array_equal(b.arrays["t"][inlet], a.arrays["t"][inlet])


def test_melt_surface_temperature_is_kept_unless_in_run(tmp_path):
    """T at melt-surface cells (IBCELL 4) is not reset from the file unless
    in_run == 1 (C/Initsv.f90:258)."""
    a = _developed(2)
    ib = a.arrays["ibcell"]
    surf = np.zeros(ib.shape, bool)
    surf[1::2, 1::2, 1::2] = ib[1::2, 1::2, 1::2] == 4
    if not surf.any():
        k0 = 1
        surf[1:-1:2, 1:-1:2, k0] = True
        ib[surf] = 4
    path = tmp_path / "rg0001c.d"
    rs.save_rg(a, path)
    b = a.clone()
    b.arrays["t"][surf] = 0.123
    rs.load_rg(b, path, in_run=0)
    assert np.all(b.arrays["t"][surf] == 0.123)
    rs.load_rg(b, path, in_run=1)
    np.testing.assert_array_equal(b.arrays["t"][surf], a.arrays["t"][surf])


# ---------------- minor-species record rs####c.d ----------------
def _case_dims(d):
    return _gd_dims(glob.glob(os.path.join(d, "gd*c.dat"))[0])


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference data files not mounted")
def test_every_reference_minor_species_file_round_trips():
    """Every rs####c.d the reference ships (17 files, two lattices), parsed
    with the lattice of its gd file and written back: the same bytes."""
    n = 0
    for f in sorted(glob.glob(os.path.join(REF, "case*", "rs*c.d"))):
        raw = open(f, "rb").read()
        rec = rs.read_rs(raw, *_case_dims(os.path.dirname(f)))
        assert rec.gfm.shape[1] == rs.NSP0 and rec.rrms.shape[-1] == rs.NRM0
        assert rs.rs_bytes(rec) == raw, f
        n += 1
    assert n == 17


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference data files not mounted")
@pytest.mark.parametrize("case", ["case0041", "case1207"])
def test_minor_species_slots_are_the_major_species(case):
    """GFM slots 1, 2, 5, 6 hold fuel, O2, CO2 and N2 (C/Initsv.f90:342-346):
    in the same run's rg and rs files they are equal on the same cells (the
    cells the minor-species solve leaves alone), and the pure-fuel cells of
    rg carry fuel 1 in rs."""
    d = os.path.join(REF, case)
    dims = _case_dims(d)
    g = rs.read_records(glob.glob(os.path.join(d, "rg*c.d"))[0], *dims)
    m = rs.read_rs(glob.glob(os.path.join(d, "rs*c.d"))[0], *dims)
    same = [m.gfm[:, slot] == g.gf[:, l - 1] for slot, l in ((0, 1), (1, 2), (4, 6), (5, 7))]
    assert np.count_nonzero(same[0]) > 1000
    for s in same[1:]:
        np.testing.assert_array_equal(s, same[0])
    pure = g.gf[:, 0] == 1.0
    assert np.count_nonzero(pure) > 0 and np.all(m.gfm[pure, 0] == 1.0)


def test_minor_species_record_rejects_corruption():
    c = cs.furnace_small(nx=6, ny=4, nz=4, seed=1)
    gfm, smf = rs.initsvm(c)
    raw = rs.rs_bytes(rs.rs_records_from(gfm, smf, itr_ms=7))
    rec = rs.read_rs(raw, *c.shape)
    assert rec.itr_ms == 7 and rs.rs_bytes(rec) == raw
    bad = bytearray(raw)
    bad[0] = 65
    with pytest.raises(ValueError):
        rs.read_rs(bytes(bad), *c.shape)
    with pytest.raises(ValueError):
        rs.read_rs(raw[:-3], *c.shape)
    with pytest.raises(ValueError):
        rs.read_rs(raw, c.shape[0] + 2, c.shape[1], c.shape[2])


def test_initsvm_fresh_start_and_restart_rules():
    """INITSVM (C/Initsv.f90:329-380): a fresh start copies fuel, O2, H2O,
    CO2, N2 into slots 1, 2, 4, 5, 6 at every cell centre; a restart takes
    the file's values except at inlet cells, which take the major species
    again with slots 3 and 7 zeroed."""
    c = cs.furnace_small(nx=8, ny=6, nz=6, seed=2)
    gfm, smf = rs.initsvm(c)
    cen = (slice(1, None, 2),) * 3
    GF = c.arrays["gf"]
    for s, l in ((0, cs.IYF), (1, cs.IYO2), (3, cs.IYH2O), (4, cs.IYCO2), (5, cs.IYN2)):
        np.testing.assert_array_equal(gfm[cen + (s,)], GF[cen + (l - 1,)])
    assert np.all(gfm[cen + (2,)] == 0.0) and np.all(gfm[cen + (6,)] == 0.0) and np.all(smf == 0.0)
    rng = np.random.default_rng(3)
    filed = rng.uniform(0.0, 1.0, gfm.shape)
    soot = rng.uniform(0.0, 1e-3, smf.shape)
    rec = rs.read_rs(rs.rs_bytes(rs.rs_records_from(filed, soot, itr_ms=12)), *c.shape)
    g2, s2 = rs.initsvm(c, rec)
    ib = c.arrays["ibcell"]
    inl = np.zeros(ib.shape, dtype=bool)
    inl[cen] = ib[cen] == 2
    opn = np.zeros(ib.shape, dtype=bool)
    opn[cen] = ib[cen] != 2
    assert inl.any()
    np.testing.assert_array_equal(g2[opn], filed[opn])
    np.testing.assert_array_equal(s2[cen], soot[cen])
    np.testing.assert_array_equal(g2[inl][:, 0], GF[..., cs.IYF - 1][inl])
    np.testing.assert_array_equal(g2[inl][:, 3], GF[..., cs.IYH2O - 1][inl])
    assert np.all(g2[inl][:, 2] == 0.0) and np.all(g2[inl][:, 6] == 0.0)


def test_committed_minor_species_record_round_trips():
    """case0041's rs0041c.d (committed compressed, with its SHA-256): parsed
    and written back byte for byte, and its fuel slot equal to rg0041c.d's
    fuel on the cells the minor-species solve leaves alone."""
    meta = json.load(open(os.path.join(GOLD, "restart_fixtures.json")))
    raw = gzip.open(os.path.join(GOLD, "rs0041c.d.gz")).read()
    assert hashlib.sha256(raw).hexdigest() == meta["rs0041c.d"]["sha256"]
    m = meta["rs0041c.d"]
    rec = rs.read_rs(raw, m["mp"], m["np"], m["lp"])
    assert rs.rs_bytes(rec) == raw
    assert rec.itr_ms == 38700 and rec.rrms.shape == (20, 24, 18, 5)
    g_meta, g_raw = _rg0041()
    g = rs.read_records(g_raw, g_meta["mp"], g_meta["np"], g_meta["lp"], g_meta["lend"])
    assert np.count_nonzero(rec.gfm[:, 0] == g.gf[:, 0]) == 2895
