"""Writes the restart and log fixtures of tests/test_restart.py and
tests/test_plt.py from data files the reference holds (run in a container
where the reference is mounted; the outputs are committed):

* rg0041c.d.gz -- combustion/case0041/rg0041c.d, the restart record of the
  TC21 case (lattice 40x48x36 from gd0041c.dat), gzip-compressed, with its
  SHA-256 in restart_fixtures.json; rs0041c.d.gz the same run's
  minor-species record;
* plt/<name>.plt -- the first lines of the case's convergence logs
  (convg, gresid, gresid_xtra, Tave of case0041), CRLF line ends kept;
* grid/gd0041c.dat, grid/gd1206c.dat and the sbc files beside them -- two
  cases (the TC21 furnace and the 154x28x30 case), with the cell count,
  melt-surface area and inflow rates their runs logged (summary####c.txt
  lines 3-10) in grid_fixtures.json.
"""
import gzip
import hashlib
import json
import os
import sys

REF = "/root/reference/GFM_Source_Package_2015-08/combustion"
HERE = os.path.dirname(os.path.abspath(__file__))
PLT_LINES = 60


def main(ref=REF):
    case = os.path.join(ref, "case0041")
    raw = open(os.path.join(case, "rg0041c.d"), "rb").read()
    with gzip.GzipFile(os.path.join(HERE, "rg0041c.d.gz"), "wb", compresslevel=9, mtime=0) as f:
        f.write(raw)
    hdr = open(os.path.join(case, "gd0041c.dat")).read().splitlines()
    mp, np_, lp = (int(v) for v in hdr[2].split()[1:4])
    meta = {"rg0041c.d": {"mp": mp, "np": np_, "lp": lp, "lend": 7, "bytes": len(raw),
                          "sha256": hashlib.sha256(raw).hexdigest()}}
    raw_rs = open(os.path.join(case, "rs0041c.d"), "rb").read()
    with gzip.GzipFile(os.path.join(HERE, "rs0041c.d.gz"), "wb", compresslevel=9, mtime=0) as f:
        f.write(raw_rs)
    meta["rs0041c.d"] = {"mp": mp, "np": np_, "lp": lp, "bytes": len(raw_rs),
                         "sha256": hashlib.sha256(raw_rs).hexdigest()}
    os.makedirs(os.path.join(HERE, "plt"), exist_ok=True)
    for name in ("convg0041c", "gresid0041c", "gresid_xtra0041c", "Tave0041c"):
        lines = open(os.path.join(case, name + ".plt"), "rb").read().split(b"\n")[:PLT_LINES]
        open(os.path.join(HERE, "plt", name + ".plt"), "wb").write(b"\n".join(lines) + b"\n")
    json.dump(meta, open(os.path.join(HERE, "restart_fixtures.json"), "w"), indent=1)
    print(meta)
    grid = {}
    os.makedirs(os.path.join(HERE, "grid"), exist_ok=True)
    for cn in ("0041", "1206"):
        d = os.path.join(ref, "case" + cn)
        for kind in ("gd", "sbc"):
            text = open(os.path.join(d, f"{kind}{cn}c.dat"), "rb").read()
            open(os.path.join(HERE, "grid", f"{kind}{cn}c.dat"), "wb").write(text)
        lines = open(os.path.join(d, f"summary{cn}c.txt")).read().splitlines()
        logged = {}
        for ln in lines[:12]:
            for key, name in (("Fuel in flow rate", "fr_f"), ("Oxygen in flow rate", "fr_ox"),
                              ("Nitrogen in flow rate", "fr_n2"), ("Heat value in with fuel", "heat_in")):
                if key in ln:
                    logged[name] = float(ln.split(":")[1].split()[0])
        for ln in lines:
            if "Energy in with flow" in ln:
                logged["q_h_in"] = float(ln.split(":")[1].split()[0])
                break
        grid[f"gd{cn}c.dat"] = dict(ncells=int(lines[2].split(":")[1]),
                                    area_melt_surf=float(lines[3].split(":")[1].split()[0]), **logged)
    json.dump(grid, open(os.path.join(HERE, "grid_fixtures.json"), "w"), indent=1)
    print(grid)


if __name__ == "__main__":
    main(*sys.argv[1:])
