"""Grid files gd####c.dat in the reference's format.

GRID3 (C/Grid3.f90:8-300) reads two title lines and the line
``r0 MP NP LP surf_length_c surf_width_c NDP0 NPT0 NSP0 NR0 [grid_type]``,
then per direction a title and pairs ``index coordinate``: the first pair
fills every lattice index up to it, each later pair (indices must rise)
fills the indices in between by adding a constant step to the previous
value, up to MP-1 (NP-1, LP-1); coordinates are divided by r0 and the last
index copies its neighbour; DX(I) = X(I+1) - X(I-1).  The cell map follows
("Computational and blocked cells"): blocks headed by the last K of the
block, rows ``I4, 2X, NP/2 digits`` (cell codes at J = 2..NP step 2 for the
I range up to the row's I; digits above 5 mean code - 10: 9 -> -1, 8 -> -2,
the half-blocked cells), then the &GRD namelist (BV_F, BA_F, BD_F).
Codes are then propagated from each blocked centre to its 26 lattice
neighbours (walls sticky) and NCELLS counts the open centres; the
melt-surface area sums DX*DR over the cells with code 4 on the K = 2
layer.  GRID3 does this once per run and logs NCELLS and the melt area in
the summary file (C/Utility.f90:99-100).

write_gd writes a case's lattice in the same format (every face listed),
so a furnace built here can be opened by the reference's tools.
"""
import re
from dataclasses import dataclass, field

import numpy as np

from .case import propagate_ibcell


@dataclass
class GridFile:
    titles: list
    header: list            # r0 MP NP LP surf_length_c surf_width_c NDP0 NPT0 NSP0 NR0 [...]
    x: np.ndarray           # lattice coordinates X(1..MP), /r0 (0-based arrays)
    dx: np.ndarray
    r: np.ndarray
    dr: np.ndarray
    z: np.ndarray
    dz: np.ndarray
    centre_codes: np.ndarray  # (MP/2, NP/2, LP/2) codes read from the map
    ibcell: np.ndarray        # (MP, NP, LP) int16, Fortran order, after propagation
    ncells: int
    area_melt_surf: float     # m^2 (area0 = r0^2 = 1)
    grd: dict = field(default_factory=dict)  # &GRD namelist values as given

    @property
    def shape(self):
        return self.ibcell.shape


def _coords(lines, pos, m, r0):
    """The coordinate pairs of one direction, C/Grid3.f90:78-98 (the same
    code reads R and Z): returns the 1-based array X(0..m) divided by r0 up
    to m-1, and the next line position."""
    pos += 1  # title
    tok = lines[pos].split()
    pos += 1
    i1, xx = int(tok[0]), float(tok[1])
    x = np.zeros(m + 1)
    x[1:i1 + 1] = xx
    while True:
        tok = lines[pos].split()
        pos += 1
        i2, xx = int(tok[0]), float(tok[1])
        if i2 <= i1:
            continue
        step = (xx - x[i1]) / (i2 - i1)
        if i2 > m or step <= 0.0:
            raise ValueError(f"grid coordinate at index {i2} out of range or not increasing")
        for i in range(i1 + 1, i2 + 1):
            x[i] = x[i - 1] + step
        i1 = i2
        if i2 >= m - 1:
            break
    x[1:m] = x[1:m] / r0
    return x, pos


def _tail(x, m, r0, full):
    """X(M) = X(M-1) and D(I) = X(I+1) - X(I-1) (C/Grid3.f90:99-104,
    133-141); with fewer than 6 lattice points in y or z (full = False) the
    one-cell branch instead, whose index 2 is divided by r0 once more
    (C/Grid3.f90:142-150).  Returns 0-based (coords, steps)."""
    d = np.zeros(m + 1)
    if full:
        x[m] = x[m - 1]
        for i in range(2, m):
            d[i] = x[i + 1] - x[i - 1]
    else:
        x[2] = x[2] / r0
        d[2] = x[2] - x[1]
    return x[1:].copy(), d[1:].copy()


def _namelist(text):
    """Values of a simple &GRD ... / namelist (name=v1,v2,...)."""
    out = {}
    body = text.split("&GRD", 1)[1].split("/", 1)[0] if "&GRD" in text.upper() else ""
    for name, vals in re.findall(r"(\w+(?:\(\s*\d+(?:\s*,\s*\d+)*\s*\))?)\s*=\s*([^=]+?)(?=\s*\w+(?:\([^)]*\))?\s*=|$)",
                                 body, flags=re.S):
        out[name.upper().replace(" ", "")] = [float(v) for v in re.split(r"[,\s]+", vals.strip()) if v]
    return out


def grd_factors(grd):
    """BV_F(1:2) and BA_F(1:3,1:2) (column-major, 6 values) of a parsed
    &GRD namelist, 1 where it gives none (C/Grid3.f90:252-255): whole arrays
    as lists, or single elements as NAME(i) / NAME(i,j)."""
    bv, ba = [1.0, 1.0], [1.0] * 6
    for key, vals in grd.items():
        m = re.fullmatch(r"(BV_F|BA_F)(?:\((\d+)(?:,(\d+))?\))?", key)
        if not m:
            continue
        dst = bv if m.group(1) == "BV_F" else ba
        if m.group(2) is None:
            start = 0
        elif m.group(1) == "BV_F":
            start = int(m.group(2)) - 1
        else:
            start = (int(m.group(2)) - 1) + 3 * (int(m.group(3) or 1) - 1)
        for q, v in enumerate(vals):
            dst[start + q] = v
    return tuple(bv), tuple(ba)


def read_gd(path_or_text):
    """Parse a gd file (path, or its text) as GRID3 does."""
    text = path_or_text
    if "\n" not in path_or_text:
        text = open(path_or_text).read()
    lines = [ln.rstrip("\r") for ln in text.split("\n")]
    titles = [lines[0], lines[1]]
    header = lines[2].split()
    r0 = float(header[0])
    mp, np_, lp = int(header[1]), int(header[2]), int(header[3])
    pos = 3
    x, pos = _coords(lines, pos, mp, r0)
    x, dx = _tail(x, mp, r0, True)
    r, pos = _coords(lines, pos, np_, r0)
    r, dr = _tail(r, np_, r0, np_ >= 6)
    z, pos = _coords(lines, pos, lp, r0)
    z, dz = _tail(z, lp, r0, lp >= 6)
    pos += 1  # "Computational and blocked cells"
    codes = np.zeros((mp // 2, np_ // 2, lp // 2), dtype=np.int16)
    k1 = 0
    while k1 < lp:
        k2 = int(lines[pos].split()[0])
        pos += 1
        i1 = 0
        while i1 < mp:
            ln = lines[pos]
            pos += 1
            i2 = int(ln[:4])
            digits = ln[6:6 + np_ // 2]
            if len(digits) != np_ // 2 or not digits.isdigit():
                raise ValueError(f"cell-map row for I = {i2} does not hold {np_ // 2} digits")
            row = np.array([int(ch) for ch in digits], dtype=np.int16)
            row[row > 5] -= 10
            codes[i1 // 2:i2 // 2, :, k1 // 2:k2 // 2] = row[None, :, None]
            i1 = i2
        k1 = k2
    grd = _namelist("\n".join(lines[pos:]))
    ib = np.zeros((mp, np_, lp), dtype=np.int16, order="F")
    ib[1::2, 1::2, 1::2] = codes
    ncells = propagate_ibcell(ib)
    melt = ib[1::2, 1::2, 1] == 4
    dxc, drc = dx[1::2], dr[1::2]
    area = float(np.sum((dxc[:, None] * drc[None, :])[melt]))
    return GridFile(titles, header, x, dx, r, dr, z, dz, codes, ib, ncells, area, grd)


def _coord_lines(title, x):
    """Every odd lattice index (the faces) from index 3 on, with X(1..3)
    equal, as GRID3 reads them back."""
    out = [title, f"{3:4d}   {float(x[2])!r}"]
    for i in range(5, len(x), 2):
        out.append(f"{i:4d}   {float(x[i - 1])!r}")
    return out


def write_gd(path, case, title="Combustion Space", surf=(0.0, 0.0)):
    """Write a case's lattice (coordinates in metres, r0 = 1, and the cell
    codes at the centres) as a gd file."""
    a, k = case.arrays, case.ctrl
    mp, np_, lp = case.shape
    codes = a["ibcell"][1::2, 1::2, 1::2]
    lines = [title, "    r0, mp, np, lp, surf_length_c, surf_width_c, ndp0, npt0, nsp0, nr0, grid_type",
             f"    1.0 {mp} {np_} {lp} {surf[0]:g} {surf[1]:g} 0 0 7 {k.nr0} 1"]
    lines += _coord_lines("X grid points", a["x"])
    lines += _coord_lines("Y grid points", a["r"])
    lines += _coord_lines("Z grid points", a["z"])
    lines.append("Computational and blocked cells")
    for kc in range(lp // 2):
        lines.append(f"{2 * kc + 2:4d}")
        for ic in range(mp // 2):
            row = "".join(str(int(v) + 10 if v < 0 else int(v)) for v in codes[ic, :, kc])
            lines.append(f"{2 * ic + 2:4d}  {row}")
    bv, ba = tuple(k.bv_f), tuple(k.ba_f)
    if bv == (1.0, 1.0) and ba == (1.0,) * 6:
        lines += ["", "&GRD", "/", ""]
    else:
        lines += ["", "&GRD", " BV_F=" + ",".join(repr(float(v)) for v in bv) + ",",
                  " BA_F=" + ",".join(repr(float(v)) for v in ba), "/", ""]
    with open(path, "w") as f:
        f.write("\n".join(lines))
