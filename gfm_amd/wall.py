"""Wall radiosity exchange (C/Qwall.f90): the inputs the sbc file gives it,
the host arrays of its boundary patches, and the radiation restart record
rr####c.d that carries them between global iterations.

``Wall`` holds the control block (``abi.GfmWallCtrl``) and the patch arrays
in patch order (``b_typ`` ... ``b_t``, as prepare_for_wall_radiation_
calculation C/Qwall.f90:393-502 lists them), the view factors ``vf``
(b_cnt x b_cnt, VF(n,m) at [n-1, m-1]) and the lattice fields ``qrs``
(MP x NP, melt-surface heat flux) and ``wall_rad_t`` (MP x NP x LP, wall
temperature at wall cells).  The device runs the exchange through
``CombustionSpace.wall_setup / wallq``.
"""
import re
import struct

import numpy as np

from .abi import NEX_MAX, NWALG_MAX, WALL_F64, WALL_I16, GfmWallCtrl

T_A, W_D, W_K, H_A, H_G = 350.0, 0.31, 2.0, 5.0, 10.0  # &RADIN defaults, C/Sbc.f90:316-330
AVG_WALL_T0 = 1600.0  # mean wall temperature before the first pass, C/Qwall.f90:69
RF_AFORM, RF_AFOVER = 0.3, 0.8  # combustion/relaxfactorc.txt, read by SETUP C/Setup.f90:489-490


def default_ctrl(**over):
    """The &RADIN defaults of SBC (C/Sbc.f90:316-330), the tolerances of
    wallq (C/Qwall.f90:1036-1037), no exhausts and no wall groups."""
    c = GfmWallCtrl()
    c.maxri1, c.maxri2, c.minri2, c.preset_vf = 100, 1000, 4, 1
    c.t_a, c.w_d, c.w_k, c.h_a, c.h_g = T_A, W_D, W_K, H_A, H_G
    c.eps_w, c.eps_m, c.eps_c = 0.8, 0.95, 0.8
    c.wall_radiosity_tolerance = c.radiation_boundary_tolerance = 1.0e-10
    c.avg_wall_T = AVG_WALL_T0
    c.q_melt_req, c.rf_aform, c.rf_afover = 1.0e6, RF_AFORM, RF_AFOVER
    for k, v in over.items():
        setattr(c, k, v)
    return c


def _radin(text):
    m = re.search(r"&RADIN(.*?)/", text, re.S | re.I)
    if not m:
        raise ValueError("no &RADIN namelist")
    vals = {}
    for key, val in re.findall(r"(\w+)\s*=\s*([^,\n]+)", m.group(1)):
        vals[key.lower()] = float(val.strip().rstrip(",").replace("d", "e").replace("D", "e"))
    return vals


def inputs_from_sbc(sbc, mp, np_, lp, ibcell):
    """The wall exchange's inputs from an sbc file (sbcfile.read_sbc(...)):
    the exhausts (C/Sbc.f90:272-284), the wall property groups and the
    IWALL array they mark (SBC_B C/Sbc.f90:288-311,402-456), the &RADIN
    values (C/Sbc.f90:316-330) and the soot calibration switch of &INPUT.
    Returns (GfmWallCtrl, iwall) with iwall (MP/2, NP/2, LP/2) int16,
    Fortran order."""
    rest = [ln for ln in sbc.rest]
    pos = 0
    while not rest[pos].strip().lower().startswith("exhausts"):
        pos += 1
    pos += 1
    nex = int(rest[pos].split()[0])
    pos += 1
    if nex > NEX_MAX:
        raise ValueError(f"{nex} exhausts, at most {NEX_MAX}")
    c = default_ctrl()
    c.nex = nex
    for m in range(nex):
        box = [int(v) for v in rest[pos].split()[:6]]
        o = rest[pos + 1].split()
        pos += 2
        for q in range(6):
            c.exh_box[m][q] = box[q]
        c.exh_orient[m], c.exh_direct[m], c.exh_type[m] = int(o[0]), int(o[1]), int(o[2])
        c.exh_frac[m], c.exh_fixed[m] = float(o[3]), float(o[4])
    while not rest[pos].strip().lower().startswith("wall properties"):
        pos += 1
    pos += 1
    iwall = np.zeros((mp // 2, np_ // 2, lp // 2), dtype=np.int16, order="F")
    n = 0
    while pos < len(rest):
        tok = rest[pos].split()
        if len(tok) < 7 or not re.fullmatch(r"-?\d+", tok[0]) or int(tok[0]) <= 0:
            break
        i1, j1, k1, i2, j2, k2, iu = (int(v) for v in tok[:7])
        w = [float(v.replace("d", "e").replace("D", "e")) for v in rest[pos + 1].split()[:4]]
        pos += 2
        n += 1
        if n > NWALG_MAX:
            raise ValueError(f"more than {NWALG_MAX} wall groups")
        c.w1_d[n - 1], c.w1_k[n - 1], c.w1_ha[n - 1], c.w1_ta[n - 1] = w
        _sbc_b(ibcell, iwall, n, (i1, j1, k1, i2, j2, k2), iu, mp, np_, lp)
    c.nwal0 = n
    v = _radin("\n".join(rest))
    for key, field in (("maxri1", "maxri1"), ("maxri2", "maxri2"), ("minri2", "minri2"), ("preset_vf", "preset_vf"),
                       ("have_crown", "have_crown")):
        if key in v:
            setattr(c, field, int(v[key]))
    for key in ("t_a", "w_d", "w_k", "h_a", "h_g", "eps_w", "eps_m", "eps_c", "height_to_ceiling"):
        if key in v:
            setattr(c, key, v[key])
    c.isoot_cal = int(sbc.number("ISOOT_CAL", 0))
    if "AFORM" in sbc.inputs:
        c.aform = sbc.number("AFORM")
    return c, iwall


def _sbc_b(ibcell, iwall, n, box, iu, mp, np_, lp):
    """SBC_B (C/Sbc.f90:402-456): along each line of the block in direction
    IU, the first wall cell from the block's start gets group n."""
    i1, j1, k1, i2, j2, k2 = box
    a = abs(iu)
    if a == 1:
        end, st = (2, -2) if iu == 1 else (mp, 2)
        for J in range(j1, j2 + 1, 2):
            for K in range(k1, k2 + 1, 2):
                for I in range(i1, end + (1 if st > 0 else -1), st):
                    if ibcell[I - 1, J - 1, K - 1] == 1:
                        iwall[I // 2 - 1, J // 2 - 1, K // 2 - 1] = n
                        break
    elif a == 2:
        end, st = (2, -2) if iu == 2 else (np_, 2)
        for I in range(i1, i2 + 1, 2):
            for K in range(k1, k2 + 1, 2):
                for J in range(j1, end + (1 if st > 0 else -1), st):
                    if ibcell[I - 1, J - 1, K - 1] == 1:
                        iwall[I // 2 - 1, J // 2 - 1, K // 2 - 1] = n
                        break
    elif a == 3:
        end, st = (2, -2) if iu == 3 else (lp, 2)
        for I in range(i1, i2 + 1, 2):
            for J in range(j1, j2 + 1, 2):
                for K in range(k1, end + (1 if st > 0 else -1), st):
                    if ibcell[I - 1, J - 1, K - 1] == 1:
                        iwall[I // 2 - 1, J // 2 - 1, K // 2 - 1] = n
                        break


class Wall:
    """Control block + host arrays of one case's wall exchange."""

    def __init__(self, case, ctrl=None, log_cap=4096):
        mp, np_, lp = case.shape
        self.shape = (mp, np_, lp)
        self.ctrl = ctrl if ctrl is not None else default_ctrl()
        self.arrays = {"qrs": np.zeros((mp, np_), order="F"), "wall_rad_t": np.zeros((mp, np_, lp), order="F"),
                       "conv_log": np.zeros(log_cap)}

    def allocate(self, b_cnt):
        """Patch arrays for b_cnt patches (after counting)."""
        self.ctrl.b_cnt = b_cnt
        for k in WALL_I16:
            self.arrays[k] = np.zeros(b_cnt, dtype=np.int16)
        for k in WALL_F64:
            self.arrays[k] = np.zeros(b_cnt)
        if self.ctrl.preset_vf == 1:
            self.arrays["vf"] = np.zeros((b_cnt, b_cnt), order="F")
            self.arrays["vf_sum"] = np.zeros(b_cnt)

    def copy(self):
        o = Wall.__new__(Wall)
        o.shape = self.shape
        o.ctrl = self.ctrl.copy()
        o.arrays = {k: np.array(v, order="F", copy=True) for k, v in self.arrays.items()}
        return o

    def qconv_face(self):
        """q_conv of each wall and melt patch at its face position on the
        lattice, 0 elsewhere: the enthalpy sink of C/Source.f90:473-509."""
        a = self.arrays
        f = np.zeros(self.shape, order="F")
        sel = (a["b_typ"] == 1) | (a["b_typ"] == 4)
        f[a["b_i"][sel] - 1, a["b_j"][sel] - 1, a["b_k"][sel] - 1] = a["q_conv"][sel]
        return f

    def log(self):
        """The mean relative radiosity change of each sweep of the last call
        (the conv_wall####c.plt lines, C/Qwall.f90:1092)."""
        return self.arrays["conv_log"][:min(self.ctrl.sweeps, self.arrays["conv_log"].size)].copy()


# ----------------------------------------------------------------------
# Radiation restart record rr####c.d
# ----------------------------------------------------------------------
# Written at the end of RAD_EMIS (C/Rad_emis.f90:290-344), read by
# rad_emis_rr (C/Rad_emis.f90:17-99): a Fortran sequential unformatted file
# (little-endian, 4-byte record lengths) of QE, QA, QL (MZ,NZ,LZ),
# wall_rad_T (MP,NP,LP), five records of sums, the patch count and
# totals, then the patch arrays (integer(2) and real(8)), the view factors
# when preset_vf = 1.
RR_SUMS = (("qeg0", "qeg_g", "qeg_w", "qeg_s", "qeg_io", "qes0"),
           ("qew0", "qew_g", "qew_w", "qew_s", "qew_io"),
           ("qes0", "svf_r", "qa_s0", "qa_s1", "qls_g", "qls_a"),
           ("q_melt", "q_wall_loss"),
           ("q_exhaust", "q_rad_inlet", "q_rad_outlet"))
RR_I4 = ("b_cnt", "preset_vf", "iterw", "iwall_count", "imelt_count", "inlet_count", "iexit_count")
RR_R8 = ("wall_area_total", "melt_area_total", "inlet_area_total", "exit_area_total", "boundary_area_total",
         "q_incident_total", "q_incident_wall", "q_incident_inlet", "q_incident_exit", "q_incident_melt",
         "qconv_total")
RR_PATCH = (("b_typ", "h"), ("b_i", "h"), ("b_j", "h"), ("b_k", "h"), ("b_orient", "h"), ("b_direct", "h"),
            ("patch_area", "d"), ("b_face_area", "d"))
RR_PATCH2 = (("radiosity", "d"), ("emis", "d"), ("q_amb", "d"), ("q_conv", "d"), ("qc", "d"), ("qeb", "d"),
             ("b_t", "d"), ("bc_i", "h"), ("bc_j", "h"), ("bc_k", "h"), ("b_exh", "h"))


def _rec(payload):
    return struct.pack("<i", len(payload)) + payload + struct.pack("<i", len(payload))


def rr_bytes(qe, qa, ql, wall, sums):
    """The bytes RAD_EMIS writes to rr####c.d.  qe, qa, ql: (MZ,NZ,LZ);
    wall: a Wall after a pass; sums: the scalars of RR_SUMS not held by the
    wall control block (qeg*, qew*, qes0, svf_r, qls_*, q_exhaust, aoxid);
    absent names are 0."""
    c, a = wall.ctrl, wall.arrays
    val = lambda k: float(getattr(c, k)) if hasattr(c, k) else float(sums.get(k, 0.0))
    out = [_rec(np.asfortranarray(x, dtype="<f8").tobytes(order="F")) for x in (qe, qa, ql, a["wall_rad_t"])]
    for grp in RR_SUMS:
        out.append(_rec(struct.pack("<" + "d" * len(grp), *(val(k) for k in grp))))
    for k in RR_I4:
        out.append(_rec(struct.pack("<i", int(getattr(c, k)))))
    for k in RR_R8:
        out.append(_rec(struct.pack("<d", float(getattr(c, k)))))
    out.append(_rec(struct.pack("<dd", float(sums.get("aoxid", 0.0)), float(c.aform))))
    out.append(_rec(struct.pack("<d", float(c.avg_wall_T))))
    for k, t in RR_PATCH:
        out.append(_rec(np.asarray(a[k], dtype="<i2" if t == "h" else "<f8").tobytes()))
    if c.preset_vf == 1:
        out.append(_rec(np.asfortranarray(a["vf"], dtype="<f8").tobytes(order="F")))
        out.append(_rec(np.asarray(a["vf_sum"], dtype="<f8").tobytes()))
    for k, t in RR_PATCH2:
        out.append(_rec(np.asarray(a[k], dtype="<i2" if t == "h" else "<f8").tobytes()))
    return b"".join(out)


def write_rr(path, qe, qa, ql, wall, sums):
    with open(path, "wb") as f:
        f.write(rr_bytes(qe, qa, ql, wall, sums))


def read_rr(src, case, ctrl=None):
    """rad_emis_rr (C/Rad_emis.f90:17-99): returns (qe, qa, ql, wall, sums)
    from the bytes or path of an rr####c.d record for the case's lattice.
    The first sums record is read as its five leading values, as the
    reference reads it."""
    buf = src if isinstance(src, (bytes, bytearray)) else open(src, "rb").read()
    mp, np_, lp = case.shape
    off = 0

    def rec():
        nonlocal off
        n = struct.unpack_from("<i", buf, off)[0]
        body = buf[off + 4:off + 4 + n]
        if struct.unpack_from("<i", buf, off + 4 + n)[0] != n:
            raise ValueError("record length markers disagree")
        off += n + 8
        return body

    cz = (mp // 2, np_ // 2, lp // 2)
    arr = lambda b, shape: np.frombuffer(b, dtype="<f8").reshape(shape, order="F").astype(np.float64).copy(order="F")
    qe, qa, ql = (arr(rec(), cz) for _ in range(3))
    wall = Wall(case, ctrl)
    wall.arrays["wall_rad_t"] = arr(rec(), (mp, np_, lp))
    sums = {}
    for q, grp in enumerate(RR_SUMS):
        b = rec()
        keys = grp[:5] if q == 0 else grp
        for k, v in zip(keys, struct.unpack_from("<" + "d" * len(keys), b)):
            sums[k] = v
    c = wall.ctrl
    for k in RR_I4:
        setattr(c, k, struct.unpack("<i", rec())[0])
    for k in RR_R8:
        setattr(c, k, struct.unpack("<d", rec())[0])
    sums["aoxid"], c.aform = struct.unpack("<dd", rec())
    c.avg_wall_T = struct.unpack("<d", rec())[0]
    for k in ("q_melt", "q_wall_loss", "q_rad_inlet", "q_rad_outlet", "qa_s0", "qa_s1"):
        setattr(c, k, sums[k])
    nb = c.b_cnt
    wall.allocate(nb)
    for k, t in RR_PATCH:
        wall.arrays[k][:] = np.frombuffer(rec(), dtype="<i2" if t == "h" else "<f8")
    if c.preset_vf == 1:
        wall.arrays["vf"][...] = np.frombuffer(rec(), dtype="<f8").reshape((nb, nb), order="F")
        wall.arrays["vf_sum"][:] = np.frombuffer(rec(), dtype="<f8")
    for k, t in RR_PATCH2:
        wall.arrays[k][:] = np.frombuffer(rec(), dtype="<i2" if t == "h" else "<f8")
    if off != len(buf):
        raise ValueError(f"{len(buf) - off} bytes after the last record")
    c.i1st_rad_done = 1
    return qe, qa, ql, wall, sums
