"""Boundary-condition files sbc####c.dat in the reference's format: the
&INPUT namelist, the pressure-reference cell and the inlet blocks that SBC
reads (C/Sbc.f90:19-70), and the inflow rates SBC forms from them
(C/Sbc.f90:90-170) and the run logs in its summary file
(C/Utility.f90:101-105: fuel, O2, N2 and total inflow, fuel heat value).

An inlet block is two lines, ``I1 J1 K1 I2 J2 K2 IU`` (lattice box, step 2,
and normal direction) and ``T YF YOX YN2 YCO2 YH2O YCK U V W`` (SI); a row
whose I1 is not positive ends the list.  Each inlet cell (cell code 2 in
the grid, counted once even where blocks overlap) carries the mass flow
theta*rho*|u_n|*A with rho from its temperature and composition at the
reference pressure (DNST_CALC after CONC, which makes H2O up to 1) and
A = AREA_C of the cell in the normal direction; the velocity passes
through the nondimensional UG0 = GMFR/DNST0 as in the reference.
"""
import re
from dataclasses import dataclass, field

import numpy as np

from .case import H_0, PG0, RU, T0, WCO2, WF, WH2O, WN2, WO2, WT0, Inlet, t_to_enth


@dataclass
class SbcFile:
    inputs: dict                 # &INPUT name -> value text (upper-case names)
    p_ref_cell: tuple            # (I_REF, J_REF, K_REF)
    p_ref: float                 # Pa, relative to PG0
    inlets: list                 # gfm_amd.case.Inlet per block
    rest: list = field(default_factory=list)  # the lines after the inlet list (exhausts, walls, &RADIN)

    def number(self, name, default=None):
        v = self.inputs.get(name.upper())
        return default if v is None else float(v.replace("d", "e").replace("D", "E"))


def _namelist(body):
    out = {}
    for tok in re.split(r"[,\n]", body):
        if "=" in tok:
            k, v = tok.split("=", 1)
            out[k.strip().upper()] = v.strip()
    return out


def read_sbc(path_or_text):
    text = path_or_text if "\n" in path_or_text else open(path_or_text).read()
    text = text.replace("\r", "")
    m = re.search(r"&INPUT(.*?)\n\s*/", text, flags=re.S | re.I)
    if not m:
        raise ValueError("no &INPUT namelist")
    inputs = _namelist(m.group(1))
    lines = text[m.end():].split("\n")
    pos = 1
    while not lines[pos].strip():
        pos += 1
    tok = lines[pos].split()
    pos += 1
    cell = (int(tok[0]), int(tok[1]), int(tok[2]))
    p_ref = float(tok[3])
    inlets = []
    while True:
        b = [int(v) for v in lines[pos].split()[:7]]
        pos += 1
        if b[0] <= 0:
            break
        s = [float(v) for v in lines[pos].split()[:10]]
        pos += 1
        t, yf, yox, yn2, yco2, yh2o, _yck, u, v, w = s
        inlets.append(Inlet(box=tuple(b[:6]), iu=b[6], T=t, yf=yf, yo2=yox, yn2=yn2, yco2=yco2, yh2o=yh2o,
                            uvw=(u, v, w)))
    return SbcFile(inputs, cell, p_ref, inlets, lines[pos:])


def inflow_rates(grid, sbc):
    """Inflow of the inlet cells of a grid (gfm_amd.gridfile.GridFile) and
    an sbc file as SBC forms it: fuel, O2 and N2 mass flow (kg/s), the fuel
    heat value fr_f*Q0 (W) and the enthalpy inflow q_h_in (W), the sum of
    T_to_enth of each inlet cell's state times its mass flow
    (C/Sbc.f90:159-165, T_to_enth C/Prop.f90:525-565)."""
    ib = grid.ibcell
    taken = np.zeros(ib.shape, dtype=bool)
    gmfr = sbc.number("GMFR")
    q0 = sbc.number("Q0", 5.0e7)
    dnst0 = PG0 * WT0 / RU / T0
    ug0 = gmfr / dnst0 / 1.0  # AREA0 = r0**2 = 1
    p = sbc.p_ref / PG0
    fr_f = fr_ox = fr_n2 = q_h_in = 0.0
    for inl in sbc.inlets:
        i1, j1, k1, i2, j2, k2 = inl.box
        iu = abs(inl.iu)
        un = inl.uvw[iu - 1] / ug0
        t = inl.T / T0
        for I in range(i1, i2 + 1, 2):
            for J in range(j1, j2 + 1, 2):
                for K in range(k1, k2 + 1, 2):
                    q = (I - 1, J - 1, K - 1)
                    if ib[q] != 2 or taken[q]:
                        continue
                    taken[q] = True
                    yf, yo, yc, yn = inl.yf, inl.yo2, inl.yco2, inl.yn2
                    s = yf + yo + yc + yn
                    yh = inl.yh2o
                    if s > 1.0:  # CONC, C/Prop.f90:28-59
                        s2 = s + yh
                        yf, yo, yc, yn, yh = yf / s2, yo / s2, yc / s2, yn / s2, yh / s2
                    else:
                        yh = 1.0 - s
                    wt = 1.0 / (yf / WF + yc / WCO2 + yh / WH2O + yo / WO2 + yn / WN2)
                    dn = max((1.0 + p) / t * wt / WT0, 1e-20)
                    if iu == 1:
                        area = grid.dr[J - 1] * grid.dz[K - 1]
                    elif iu == 2:
                        area = grid.dx[I - 1] * grid.dz[K - 1]
                    else:
                        area = grid.dx[I - 1] * grid.dr[J - 1]
                    g0 = 1.0 * dn * dnst0 * abs(un) * ug0 * area
                    fr_f += g0 * yf
                    fr_ox += g0 * yo
                    fr_n2 += g0 * yn
                    h = float(t_to_enth([np.array(v) for v in (yf, yo, yc, yn, yh)], np.array(t)))
                    q_h_in += h * H_0 * g0
    return dict(fr_f=fr_f, fr_ox=fr_ox, fr_n2=fr_n2, heat_in=fr_f * q0, q_h_in=q_h_in)
