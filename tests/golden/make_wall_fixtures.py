"""Writes wall_fixtures.json, the results of the wall radiosity exchange
that two reference runs logged, for tests/test_oracle_wall.py (run in a
container where the reference is mounted; the output is committed).

The run's last global iteration ran its radiation pass -- volume part,
then wall_q_amb, wallq, QRN_PRN (C/Qrn.f90:218-221) -- on the state its
records hold (C/A0main.f90:204-233), so these files are that pass's
results:

* it####m.dat (QRN_PRN, C/Qrn.f90:371-394): the mean wall temperature,
  the largest gas temperature, the heat to the melt q_melt (full
  precision), and the melt-surface heat flux QRS on the (I, J) cells
  (E11.4);
* twall####c.txt (wallq, C/Qwall.f90:1323-1337): the interior wall patch
  count, the patch count and the wall temperature wall_rad_T of each wall
  cell (pform, 7 digits);
* Tave####c.plt (GFLOW, C/Gflow.f90:399): the mean wall temperature of the
  previous pass on the last gas iteration's line -- an input, since exits
  of exh_type 0 take a fraction of it during the exchange;
* soot_cal####c.plt (C/Qwall.f90:1319) and it####t.dat (INITSV,
  C/Initsv.f90:428): with isoot_cal = 1 (case1206), the soot formation
  constant before and after the pass and the melter's demand q_melt_req.
"""
import json
import os
import sys

REF = "/root/reference/GFM_Source_Package_2015-08/combustion"
HERE = os.path.dirname(os.path.abspath(__file__))


def parse_pform(lines):
    """(I, J, K, value) of every nonzero entry of a pform dump: per Z(K)
    block, column groups headed 'X  /  Y', rows I = 2, 4, ..."""
    out, k, js, ii, jcol = [], None, [], 2, 2
    for ln in lines:
        s = ln.strip()
        if s.startswith("Z("):
            k, jcol = int(s[2:s.index(")")]), 2
            continue
        if s.startswith("X  /  Y"):
            n = len(s.split()) - 3
            js, jcol, ii = list(range(jcol, jcol + 2 * n, 2)), jcol + 2 * n, 2
            continue
        v = s.split()
        if k is None or len(v) != len(js) + 1:
            continue
        for q, j in enumerate(js):
            x = float(v[q + 1])
            if x != 0.0:
                out.append([ii, j, k, x])
        ii += 2
    return out


def parse_qrs(lines):
    """(I, J, QRS) of every nonzero entry of the flux table of it####m.dat
    (QRN_PRN writes it with the format (F11.3, 11E11.4))."""
    out, js, ii, jcol = [], [], 2, 2
    for ln in lines:
        s = ln.strip()
        if s.startswith("X  /  Y"):
            n = len(s.split()) - 3
            js, jcol, ii = list(range(jcol, jcol + 2 * n, 2)), jcol + 2 * n, 2
            continue
        if not js or not s:
            continue
        # fixed width (F11.3, 11E11.4): a negative value fills its field
        v = [ln[11 * q:11 * q + 11] for q in range(len(js) + 1)]
        for q, j in enumerate(js):
            x = float(v[q + 1])
            if x != 0.0:
                out.append([ii, j, x])
        ii += 2
    return out


def main(ref=REF):
    meta = {}
    for cn in ("0041", "1206"):
        d = os.path.join(ref, "case" + cn)
        m = open(os.path.join(d, f"it{cn}m.dat")).read().splitlines()
        ent = {"avg_wall_T": float(m[1].split()[0]), "tmax_c": float(m[2].split()[0]),
               "q_melt": float(m[3].split()[0]), "qrs": parse_qrs(m[4:])}
        tw = open(os.path.join(d, f"twall{cn}c.txt")).read().splitlines()
        head = " ".join(tw[:3]).split()
        ent["iwall_count"] = int(head[head.index("cells:") + 1])
        ent["b_cnt"] = int(head[head.index("patches:") + 1])
        ent["wall_rad_T"] = parse_pform(tw)
        tave = open(os.path.join(d, f"Tave{cn}c.plt")).read().strip().splitlines()[-1].split()
        ent["itr_gas"] = int(tave[0])
        ent["avg_wall_T_prev"] = float(tave[3])
        sc = os.path.join(d, f"soot_cal{cn}c.plt")
        if os.path.exists(sc):
            rows = [r.split() for r in open(sc).read().splitlines() if r.strip() and not r.lstrip().startswith("#")]
            ent["aform_prev"] = float(rows[-2][2])
            ent["aform"] = float(rows[-1][2])
            ent["q_melt_req"] = float(open(os.path.join(d, f"it{cn}t.dat")).read().splitlines()[1].split()[0])
        meta["case" + cn] = ent
    json.dump(meta, open(os.path.join(HERE, "wall_fixtures.json"), "w"), separators=(",", ":"))
    for cn, e in meta.items():
        print(cn, {k: v for k, v in e.items() if not isinstance(v, list)}, len(e["wall_rad_T"]), len(e["qrs"]))


if __name__ == "__main__":
    main(*sys.argv[1:])
