"""Convergence logs of the combustion space in the reference's text format.

The reference appends one line per SIMPLER iteration to ``.plt`` files that
its GUI and RunPlot read (file catalogue D/LaTeX/Files/GFM4-Files.tex):

* ``convg####c.plt``: itr_gas, res_mass, res_mass_max, format (i7,4e25.16)
  (C/Sphase.f90:53), header C/Utility.f90:1076-1080;
* ``gresid####c.plt``: enthalpy, pressure, x/y/z-momentum residuals after
  the solve, (i7,5e24.16) (C/Gflow.f90:259-261), header
  C/Utility.f90:434-446; ``gresidp####c.plt`` the same before the solve
  (C/Gflow.f90:262-264, C/Utility.f90:352-365);
* ``gresid_xtra####c.plt``: fuel, O2, CO2, N2, then GF residuals 4 and 5
  (the header names them Epsilon, k), (i7,6e24.16) (C/Gflow.f90:267-269,
  C/Utility.f90:449-464); ``gresid_xtrap####c.plt`` before the solve;
* ``Tave####c.plt``: mass-weighted mean gas T and mean exit T, (i5,2e22.14),
  with the mean wall T as a third value once radiation has run
  (C/Gflow.f90:392-401, C/Utility.f90:272-290).

Headers written list-directed carry the leading blank of list-directed
output; lines end in CR LF as in the reference's files (newline='\\n' for
the platform's own line ends).
"""
import os

__all__ = ["fortran_e", "fortran_i", "PlotLogs", "LOGS"]


def fortran_i(v, w):
    return f"{int(v):>{w}d}"


def fortran_e(x, w, d):
    """Fortran Ew.d output of x: 0.d1..dd mantissa, E+xx exponent (3 digits
    and no E beyond 99), right-justified in w columns."""
    x = float(x)
    if x != x:
        return "NaN".rjust(w)
    if x == 0.0:
        body, e = "0." + "0" * d, 0
    else:
        s = f"{abs(x):.{d - 1}e}"
        mant, ex = s.split("e")
        body = "0." + mant.replace(".", "")
        e = int(ex) + 1
    sign = "-" if x < 0 or (x == 0.0 and str(x).startswith("-")) else ""
    exp = f"E{e:+03d}" if abs(e) <= 99 else f"{e:+04d}"
    return (sign + body + exp).rjust(w)


def _list_directed(text):
    return " " + text


def _slash_block(text):
    """'(/"..."/)': a blank line, the text, a blank line."""
    return ["", text, ""]


HEADERS = {
    "convg": [_list_directed("# Combustion space mean and max mass residual - log plot"),
              _list_directed("# PlotData"),
              _list_directed("# iteration  average mass residual  max mass residual")],
    "gresid": [_list_directed("# Mean residuals - log plot"), _list_directed("# PlotData ")] + _slash_block(
        "# Iter " + "        Enthalpy        " + "        Pressure        " + "       x-momentum       " +
        "       y-momentum       " + "       z-momentum       "),
    "gresidp": [_list_directed("# Mean residuals (pre-solve) - log plot"), _list_directed("# PlotData ")] + _slash_block(
        "# Iter " + "      Pre-Enthalpy      " + "      Pre-Pressure      " + "     Pre-x-momentum     " +
        "     Pre-y-momentum     " + "     Pre-z-momentum     "),
    "gresid_xtra": [_list_directed("# Mean residuals: major species, k-epsilon - log plot"),
                    _list_directed("# PlotData ")] + _slash_block(
        "# Iter " + "        Fuel            " + "        O2              " + "        CO2             " +
        "        N2              " + "        Epsilon         " + "        k               "),
    "gresid_xtrap": [_list_directed("# Mean residuals: major species, k-epsilon (pre-solve) - log plot"),
                     _list_directed("# PlotData ")] + _slash_block(
        "# Iter " + "      Pre-Fuel          " + "      Pre-O2            " + "      Pre-CO2           " +
        "      Pre-N2            " + "      Pre-Epsilon       " + "      Pre-k             "),
    "Tave": [_list_directed("# Combustion space mean temperature"), _list_directed("# PlotData ")] + _slash_block(
        "# Iter      " + "Mean Gas Temp (K)     " + "Mean Exit Temp        " + "Mean Wall Temp"),
}

# (integer width, real width, digits) of each log's record format
FORMATS = {"convg": (7, 25, 16), "gresid": (7, 24, 16), "gresidp": (7, 24, 16), "gresid_xtra": (7, 24, 16),
           "gresid_xtrap": (7, 24, 16), "Tave": (5, 22, 14)}
LOGS = tuple(FORMATS)


def format_record(name, itr, values):
    iw, w, d = FORMATS[name]
    return fortran_i(itr, iw) + "".join(fortran_e(v, w, d) for v in values)


def record_values(name, ctrl, avg_wall_t=None):
    """The values the reference writes for one iteration, from the ctrl
    block the library returns (C/Sphase.f90:53, C/Gflow.f90:259-269,392-401)."""
    k = ctrl
    if name == "convg":
        return [k.res_mass, k.res_mass_max]
    if name == "gresid":
        return [k.h_resid, k.p_resid, k.u_resid[0], k.u_resid[1], k.u_resid[2]]
    if name == "gresidp":
        return [k.h_resid_pre, k.p_resid_pre, k.u_resid_pre[0], k.u_resid_pre[1], k.u_resid_pre[2]]
    if name in ("gresid_xtra", "gresid_xtrap"):
        r = k.gf_resid if name == "gresid_xtra" else k.gf_resid_pre
        return [r[0], r[1], r[5], r[6], r[3], r[4]]
    if name == "Tave":
        v = [k.tave, k.tmean_ex]
        if k.i1st_rad_done == 1 and avg_wall_t is not None:
            v.append(avg_wall_t)
        return v
    raise KeyError(name)


class PlotLogs:
    """The reference's per-iteration logs of one run in a case directory
    (run number runum: files <name><runum:04d>c.plt).  A fresh run writes
    the headers; a restart (irstyp > 0) appends to existing files, as the
    reference's *_file_init routines do."""

    def __init__(self, casedir, runum, names=("convg", "gresid", "gresid_xtra", "Tave"), irstyp=0, newline="\r\n"):
        self.nl = newline
        self.files = {}
        for n in names:
            path = os.path.join(casedir, f"{n}{runum:04d}c.plt")
            append = irstyp > 0 and os.path.exists(path)
            f = open(path, "a" if append else "w", newline="")
            if not append:
                f.write("".join(h + self.nl for h in HEADERS[n]))
            self.files[n] = f

    def write(self, ctrl, avg_wall_t=None):
        """One line per log for the iteration ctrl.itr_gas."""
        for n, f in self.files.items():
            f.write(format_record(n, ctrl.itr_gas, record_values(n, ctrl, avg_wall_t)) + self.nl)

    def close(self):
        for f in self.files.values():
            f.close()
        self.files = {}

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
