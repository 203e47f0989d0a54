"""Restart records of the combustion space in the reference's own format.

``rg####c.d`` is a Fortran sequential unformatted file (little-endian, a
4-byte length before and after every record), written by SAV3F
(C/Sav3f.f90:21-88) and read by read_restart_data (C/Initsv.f90:225-316):

* per cell centre, I = 2..MP, J = 2..NP, K = 2..LP in steps of 2 (K
  innermost), one record (P, U(I+1), V(J+1), W(K+1), T, THETA) -- the face
  index wraps to MP-1 (NP-1, LP-1) on the last layer -- and one record
  GF(I,J,K,1:LEND);
* bs_hprev (the whole lattice array), T_surf_prev (MP/2 x NP/2);
* (FACTOR, FADJ), i1st_rad_qe_done, i1st_rad_done, itr_gas,
  (itr_rad, itr_gend), interval_rad (4-byte integers).

Only the single-phase layout is written (NPHAS = 1: the particle records of
C/Sav3f.f90:52-73 never occur on this path).  A device-resident run
checkpoints by downloading its state (CombustionSpace.download) and
calling save_rg; it resumes by load_rg on a freshly built case followed by
initialisation, as the reference does at every global iteration
(C/A0main.f90:157-251).
"""
import struct
from dataclasses import dataclass, field

import numpy as np

from .case import IEPS, IK, dnst_calc_at, ther_at

LEND = 7  # GF(:,:,:,1:LEND) in the restart records, C/Var_mod.f90:125
T_SURF_INIT = 1500.0  # T_surf_prev before any surface relaxation, C/Setup.f90:755-756
INTERVAL_RAD = 100  # default gas iterations between radiation passes, C/Var_mod.f90:468


@dataclass
class RestartRecords:
    """The contents of one rg####c.d file, in record order."""
    mp: int
    np: int
    lp: int
    cell6: np.ndarray   # (cells, 6): P, U(I+1), V(J+1), W(K+1), T, THETA
    gf: np.ndarray      # (cells, lend)
    bs_hprev: np.ndarray  # (mp, np, lp), Fortran order
    t_surf_prev: np.ndarray  # (mp//2, np//2), Fortran order
    factor: float = 1.0
    fadj: float = 0.0
    i1st_rad_qe_done: int = 0
    i1st_rad_done: int = 0
    itr_gas: int = 0
    itr_rad: int = 0
    itr_gend: int = 0
    interval_rad: int = INTERVAL_RAD
    extra: dict = field(default_factory=dict)

    @property
    def lend(self):
        return self.gf.shape[1]


def _centres(m):
    """0-based centre indices of the lattice I = 2..M step 2, and the index
    of the face each centre's record carries (I+1, or M-1 on the last)."""
    c = np.arange(1, m, 2)
    return c, np.where(c + 1 < m, c + 1, m - 2)


def _cell_index(mp, np_, lp):
    ci, fi = _centres(mp)
    cj, fj = _centres(np_)
    ck, fk = _centres(lp)
    I, J, K = np.meshgrid(ci, cj, ck, indexing="ij")
    FI, FJ, FK = np.meshgrid(fi, fj, fk, indexing="ij")
    return tuple(a.ravel() for a in (I, J, K, FI, FJ, FK))  # C order: K innermost, I outermost


def _put(f, payload):
    f.write(struct.pack("<i", len(payload)))
    f.write(payload)
    f.write(struct.pack("<i", len(payload)))


class _Reader:
    def __init__(self, buf):
        self.buf = buf
        self.off = 0

    def record(self):
        (n,) = struct.unpack_from("<i", self.buf, self.off)
        start = self.off + 4
        end = start + n
        if end + 4 > len(self.buf):
            raise ValueError(f"record of {n} bytes at offset {self.off} runs past the end of the file")
        (n2,) = struct.unpack_from("<i", self.buf, end)
        if n2 != n:
            raise ValueError(f"record markers {n} and {n2} disagree at offset {self.off}")
        self.off = end + 4
        return self.buf[start:end]

    def doubles(self, count):
        r = self.record()
        if len(r) != 8 * count:
            raise ValueError(f"expected a record of {count} doubles, found {len(r)} bytes")
        return np.frombuffer(r, dtype="<f8").copy()

    def ints(self, count):
        r = self.record()
        if len(r) != 4 * count:
            raise ValueError(f"expected a record of {count} integers, found {len(r)} bytes")
        return [int(v) for v in np.frombuffer(r, dtype="<i4")]


def _cell_dtype(lend):
    """One cell's pair of records: marker, 6 doubles, marker, marker, lend
    doubles, marker."""
    return np.dtype([("m0", "<i4"), ("c6", "<f8", (6,)), ("m1", "<i4"), ("m2", "<i4"), ("gf", "<f8", (lend,)),
                     ("m3", "<i4")])


def read_records(src, mp, np_, lp, lend=LEND):
    """Parse rg####c.d (a path or the file's bytes) for a lattice of
    (mp, np_, lp); raises ValueError on any record that does not match."""
    buf = src if isinstance(src, (bytes, bytearray, memoryview)) else open(src, "rb").read()
    buf = bytes(buf)
    ncell = (mp // 2) * (np_ // 2) * (lp // 2)
    dt = _cell_dtype(lend)
    if len(buf) < ncell * dt.itemsize:
        raise ValueError(f"file of {len(buf)} bytes is shorter than the {ncell} cell records")
    cells = np.frombuffer(buf, dtype=dt, count=ncell)
    for m, n in (("m0", 48), ("m1", 48), ("m2", 8 * lend), ("m3", 8 * lend)):
        bad = np.nonzero(cells[m] != n)[0]
        if len(bad):
            raise ValueError(f"cell record {bad[0]}: marker {cells[m][bad[0]]}, expected {n}")
    cell6 = cells["c6"].copy()
    gf = cells["gf"].copy()
    rd = _Reader(buf)
    rd.off = ncell * dt.itemsize
    bs_hprev = rd.doubles(mp * np_ * lp).reshape((mp, np_, lp), order="F")
    t_surf = rd.doubles((mp // 2) * (np_ // 2)).reshape((mp // 2, np_ // 2), order="F")
    factor, fadj = rd.doubles(2)
    (qe,) = rd.ints(1)
    (rad,) = rd.ints(1)
    rec = RestartRecords(mp, np_, lp, cell6, gf, bs_hprev, t_surf, float(factor), float(fadj), qe, rad)
    # itr_gas, (itr_rad, itr_gend), interval_rad: read_restart_data tolerates
    # their absence (read(..., err=380), C/Initsv.f90:297-303)
    try:
        (rec.itr_gas,) = rd.ints(1)
        rec.itr_rad, rec.itr_gend = rd.ints(2)
        (rec.interval_rad,) = rd.ints(1)
    except (ValueError, struct.error):
        rec.extra["truncated_tail"] = True
    if rd.off != len(rd.buf):
        raise ValueError(f"{len(rd.buf) - rd.off} bytes after the last record")
    return rec


def write_records(path, rec):
    """Write rec as SAV3F does (C/Sav3f.f90:27-85)."""
    with open(path, "wb") as f:
        f.write(records_bytes(rec))


def records_bytes(rec):
    import io
    f = io.BytesIO()
    cells = np.empty(rec.cell6.shape[0], dtype=_cell_dtype(rec.lend))
    cells["m0"] = cells["m1"] = 48
    cells["m2"] = cells["m3"] = 8 * rec.lend
    cells["c6"] = rec.cell6
    cells["gf"] = rec.gf
    f.write(cells.tobytes())
    _put(f, np.asarray(rec.bs_hprev, dtype="<f8").tobytes(order="F"))
    _put(f, np.asarray(rec.t_surf_prev, dtype="<f8").tobytes(order="F"))
    _put(f, np.array([rec.factor, rec.fadj], dtype="<f8").tobytes())
    _put(f, np.array([rec.i1st_rad_qe_done], dtype="<i4").tobytes())
    _put(f, np.array([rec.i1st_rad_done], dtype="<i4").tobytes())
    if not rec.extra.get("truncated_tail"):
        _put(f, np.array([rec.itr_gas], dtype="<i4").tobytes())
        _put(f, np.array([rec.itr_rad, rec.itr_gend], dtype="<i4").tobytes())
        _put(f, np.array([rec.interval_rad], dtype="<i4").tobytes())
    return f.getvalue()


def records_from_case(case, t_surf_prev=None, i1st_rad_qe_done=0, itr_rad=0, interval_rad=INTERVAL_RAD, lend=LEND):
    """The records SAV3F writes for a case's host state (C/Sav3f.f90:27-48,
    75-85)."""
    a, k = case.arrays, case.ctrl
    mp, np_, lp = case.shape
    I, J, K, FI, FJ, FK = _cell_index(mp, np_, lp)
    ug = a["ug"]
    cell6 = np.stack([a["p"][I, J, K], ug[FI, J, K, 0], ug[I, FJ, K, 1], ug[I, J, FK, 2], a["t"][I, J, K],
                      a["theta"][I, J, K]], axis=1)
    gf = np.stack([a["gf"][I, J, K, l] for l in range(lend)], axis=1)
    if t_surf_prev is None:
        t_surf_prev = np.full((mp // 2, np_ // 2), T_SURF_INIT)
    return RestartRecords(mp, np_, lp, cell6, gf, np.asfortranarray(a["bs_hprev"]), np.asfortranarray(t_surf_prev),
                          float(k.factor), float(k.fadj), int(i1st_rad_qe_done), int(k.i1st_rad_done),
                          int(k.itr_gas), int(itr_rad), int(k.itr_gend), int(interval_rad))


def save_rg(case, path, **kw):
    """SAV3F: write the case's host state to path (rg####c.d)."""
    write_records(path, records_from_case(case, **kw))


def apply_records(case, rec, in_run=0, interval_rad=INTERVAL_RAD, itr_gas_tmp=0):
    """read_restart_data (C/Initsv.f90:239-313) on a case's host state.

    Per centre: P always; at an inlet cell of a real restart (itr_gas <= 0)
    only k and epsilon; elsewhere the +faces of U, V, W unless that face is
    an inlet or wall, T unless the cell is melt surface (and in_run == 0),
    THETA, GF(1:LEND), then DNST_CALC and THER.  Then bs_hprev, FACTOR,
    FADJ, the radiation flags and -- on a real restart -- itr_gas, itr_gend
    (restarted when it lies behind itr_gas or interval_rad changed), and
    maxgi += the restored itr_gas (itr_gas_tmp on a loop-forced restart).
    Returns the values the case model does not hold: T_surf_prev,
    i1st_rad_qe_done, itr_rad."""
    a, k = case.arrays, case.ctrl
    mp, np_, lp = case.shape
    if (rec.mp, rec.np, rec.lp) != (mp, np_, lp):
        raise ValueError(f"restart lattice {(rec.mp, rec.np, rec.lp)} is not the case's {(mp, np_, lp)}")
    ib = a["ibcell"]
    I, J, K, _, _, _ = _cell_index(mp, np_, lp)
    a["p"][I, J, K] = rec.cell6[:, 0]
    real = k.itr_gas <= 0
    inlet = (ib[I, J, K] == 2) & real
    for l in (IK, IEPS):
        if l <= rec.lend:
            a["gf"][I[inlet], J[inlet], K[inlet], l - 1] = rec.gf[inlet, l - 1]
    rest = ~inlet
    Ir, Jr, Kr = I[rest], J[rest], K[rest]
    for d, (di, dj, dk) in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
        Fi, Fj, Fk = Ir + di, Jr + dj, Kr + dk
        inside = (Fi < mp) & (Fj < np_) & (Fk < lp)
        sel = np.zeros(Ir.shape, dtype=bool)
        code = ib[np.minimum(Fi, mp - 1), np.minimum(Fj, np_ - 1), np.minimum(Fk, lp - 1)]
        sel[inside] = (code[inside] != 2) & (code[inside] != 1)
        a["ug"][Fi[sel], Fj[sel], Fk[sel], d] = rec.cell6[rest, 1 + d][sel]
    tset = (ib[Ir, Jr, Kr] != 4) | (in_run == 1)
    a["t"][Ir[tset], Jr[tset], Kr[tset]] = rec.cell6[rest, 4][tset]
    a["theta"][Ir, Jr, Kr] = rec.cell6[rest, 5]
    for l in range(rec.lend):
        a["gf"][Ir, Jr, Kr, l] = rec.gf[rest, l]
    sl = (Ir, Jr, Kr)
    dnst_calc_at(a, sl)
    ther_at(a, k, sl)
    a["bs_hprev"][...] = rec.bs_hprev
    k.factor, k.fadj = rec.factor, rec.fadj
    if rec.i1st_rad_done == 1 or k.i1st_rad_done == 1:
        k.i1st_rad_done = 1
    itr_rad = 0
    if real and not rec.extra.get("truncated_tail"):
        itr_gas_tmp = rec.itr_gas
        k.itr_gas = rec.itr_gas
        itr_rad = rec.itr_rad
        k.itr_gend = rec.itr_gend
        if k.itr_gend <= k.itr_gas or rec.interval_rad != interval_rad:
            k.itr_gend = k.itr_gas + interval_rad
    k.maxgi = k.maxgi + itr_gas_tmp
    return {"t_surf_prev": rec.t_surf_prev, "i1st_rad_qe_done": rec.i1st_rad_qe_done, "itr_rad": itr_rad}


def load_rg(case, path, **kw):
    """read_restart_data: apply rg####c.d at path to the case's host state."""
    mp, np_, lp = case.shape
    return apply_records(case, read_records(path, mp, np_, lp), **kw)


# ----------------------------------------------------------------------
# Minor-species restart record rs####c.d
# ----------------------------------------------------------------------
# Written by sav3fm (C/Sav3f.f90:96-116), read by INITSVM
# (C/Initsv.f90:329-380): per cell centre, I = 2..MP, J, K in steps of 2
# (K innermost), one record GFM(I,J,K,1:NSP0), SMF(I,J,K); then the whole
# lattice arrays bs_soot_prev and sfp_soot_prev, RRMS(MP/2, NP/2, LP/2, NRM0)
# (the minor-species reaction rates, C/Setup.f90:911) and itr_ms.
NSP0 = 7  # minor species, C/Var_mod.f90:218
NRM0 = 5  # minor-species reactions of RRMS


@dataclass
class MinorSpeciesRecords:
    """The contents of one rs####c.d file, in record order."""
    mp: int
    np: int
    lp: int
    gfm: np.ndarray             # (cells, nsp0): GFM(I,J,K,1:NSP0) per cell centre, K innermost
    smf: np.ndarray             # (cells,): SMF, the soot mass fraction
    bs_soot_prev: np.ndarray    # (mp, np, lp), Fortran order
    sfp_soot_prev: np.ndarray   # (mp, np, lp), Fortran order
    rrms: np.ndarray            # (mp//2, np//2, lp//2, nrm0), Fortran order
    itr_ms: int = 0

    @property
    def nsp0(self):
        return self.gfm.shape[1]


def read_rs(src, mp, np_, lp, nsp0=NSP0, nrm0=NRM0):
    """Parse rs####c.d (a path or the file's bytes) for a lattice of
    (mp, np_, lp); raises ValueError on any record that does not match."""
    buf = src if isinstance(src, (bytes, bytearray, memoryview)) else open(src, "rb").read()
    buf = bytes(buf)
    ncell = (mp // 2) * (np_ // 2) * (lp // 2)
    n = 8 * (nsp0 + 1)
    dt = np.dtype([("m0", "<i4"), ("v", "<f8", (nsp0 + 1,)), ("m1", "<i4")])
    if len(buf) < ncell * dt.itemsize:
        raise ValueError(f"file of {len(buf)} bytes is shorter than the {ncell} cell records")
    cells = np.frombuffer(buf, dtype=dt, count=ncell)
    for m in ("m0", "m1"):
        bad = np.nonzero(cells[m] != n)[0]
        if len(bad):
            raise ValueError(f"cell record {bad[0]}: marker {cells[m][bad[0]]}, expected {n}")
    r = _Reader(buf)
    r.off = ncell * dt.itemsize
    nlat = mp * np_ * lp
    bsp = r.doubles(nlat).reshape((mp, np_, lp), order="F")
    sfp = r.doubles(nlat).reshape((mp, np_, lp), order="F")
    rr = r.doubles(ncell * nrm0).reshape((mp // 2, np_ // 2, lp // 2, nrm0), order="F")
    (itr_ms,) = r.ints(1)
    if r.off != len(buf):
        raise ValueError(f"{len(buf) - r.off} bytes after the last record")
    v = cells["v"]
    return MinorSpeciesRecords(mp, np_, lp, v[:, :nsp0].copy(), v[:, nsp0].copy(), bsp, sfp, rr, itr_ms)


def rs_bytes(rec):
    """The rs####c.d file of a MinorSpeciesRecords, as sav3fm writes it."""
    n = 8 * (rec.nsp0 + 1)
    dt = np.dtype([("m0", "<i4"), ("v", "<f8", (rec.nsp0 + 1,)), ("m1", "<i4")])
    cells = np.zeros(len(rec.smf), dtype=dt)
    cells["m0"] = n
    cells["m1"] = n
    cells["v"][:, :rec.nsp0] = rec.gfm
    cells["v"][:, rec.nsp0] = rec.smf
    out = [cells.tobytes()]
    for a in (rec.bs_soot_prev, rec.sfp_soot_prev, rec.rrms):
        p = np.asarray(a, dtype="<f8").tobytes(order="F")
        out.append(struct.pack("<i", len(p)) + p + struct.pack("<i", len(p)))
    p = struct.pack("<i", int(rec.itr_ms))
    out.append(struct.pack("<i", 4) + p + struct.pack("<i", 4))
    return b"".join(out)


def write_rs(path, rec):
    with open(path, "wb") as f:
        f.write(rs_bytes(rec))


def initsvm(case, rec=None):
    """The minor-species fields INITSVM sets up (C/Initsv.f90:329-380):
    returns (gfm, smf) on the lattice, gfm of shape (MP, NP, LP, NSP0).
    Fresh start (rec None): every cell centre takes fuel, O2, H2O, CO2 and
    N2 from the major species into slots 1, 2, 4, 5, 6, the rest zero, and
    smf = 0.  Restart: the file's values, except at inlet cells, where the
    five come from the major species again and slots 3 and 7 are zeroed."""
    from .case import IYF, IYO2, IYH2O, IYCO2, IYN2
    mp, np_, lp = case.shape
    nsp0 = NSP0 if rec is None else rec.nsp0
    gfm = np.zeros((mp, np_, lp, nsp0), order="F")
    smf = np.zeros((mp, np_, lp), order="F")
    GF = case.arrays["gf"]
    ib = case.arrays["ibcell"]
    cen = (slice(1, None, 2),) * 3
    take = ((0, IYF), (1, IYO2), (3, IYH2O), (4, IYCO2), (5, IYN2))
    if rec is None:
        for s, l in take:
            gfm[cen + (s,)] = GF[cen + (l - 1,)]
        return gfm, smf
    if (rec.mp, rec.np, rec.lp) != (mp, np_, lp):
        raise ValueError("restart record and case lattices differ")
    I, J, K, _, _, _ = _cell_index(mp, np_, lp)
    gfm[I, J, K, :] = rec.gfm
    smf[I, J, K] = rec.smf
    inl = ib[I, J, K] == 2
    for s, l in take:
        gfm[I[inl], J[inl], K[inl], s] = GF[I[inl], J[inl], K[inl], l - 1]
    gfm[I[inl], J[inl], K[inl], 2] = 0.0
    gfm[I[inl], J[inl], K[inl], 6] = 0.0
    return gfm, smf


def rs_records_from(gfm, smf, bs_soot_prev=None, sfp_soot_prev=None, rrms=None, itr_ms=0):
    """MinorSpeciesRecords of lattice fields (what sav3fm writes)."""
    mp, np_, lp, nsp0 = gfm.shape
    I, J, K, _, _, _ = _cell_index(mp, np_, lp)
    z = lambda: np.zeros((mp, np_, lp), order="F")
    return MinorSpeciesRecords(mp, np_, lp, gfm[I, J, K, :].copy(), smf[I, J, K].copy(),
                               z() if bs_soot_prev is None else bs_soot_prev,
                               z() if sfp_soot_prev is None else sfp_soot_prev,
                               np.zeros((mp // 2, np_ // 2, lp // 2, NRM0), order="F") if rrms is None else rrms,
                               itr_ms)
