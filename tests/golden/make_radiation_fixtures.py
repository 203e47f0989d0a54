"""Writes the radiation fixtures of tests/test_oracle_radiation.py from data
files the reference holds (run in a container where the reference is
mounted; the outputs are committed):

* rg1206c.d.gz, rs1206c.d.gz -- the restart and minor-species records of
  combustion/case1206 (lattice 154x28x30 from grid/gd1206c.dat),
  gzip-compressed (case0041's are written by make_restart_fixtures.py);
* radiation_fixtures.json -- for case0041 and case1206: the SHA-256 of the
  records, and the last line of the run's radiation log rad_detail####c.plt
  (C/Qrn.f90:205-217): the gas iteration, then the emission of H2O+CO2 and
  of soot (RAD0), their sum, the emission of the traced emitters, the
  absorption at the melt surface, walls, open cells, exits and inlets, the
  absorbed total and the relative difference.  The run's last global
  iteration wrote the records (SAV3F, sav3fm) and then ran this radiation
  pass on the state they hold (C/A0main.f90:204-233), so the line is the
  pass's result on that state.
"""
import gzip
import hashlib
import json
import os
import sys

REF = "/root/reference/GFM_Source_Package_2015-08/combustion"
HERE = os.path.dirname(os.path.abspath(__file__))
KEYS = ("itr_gas", "tot_qe_h2o_co2", "tot_qe_soot", "qe_sum", "qe_sum_from_open", "qa_sum_to_melt",
        "qa_sum_to_wall", "qa_sum_to_open", "qa_sum_to_exit", "qa_sum_to_inlet", "qa_all", "q_diff")


def main(ref=REF):
    meta = {}
    for cn in ("0041", "1206"):
        d = os.path.join(ref, "case" + cn)
        ent = {}
        for kind in ("rg", "rs"):
            raw = open(os.path.join(d, f"{kind}{cn}c.d"), "rb").read()
            ent[f"{kind}_sha256"] = hashlib.sha256(raw).hexdigest()
            if cn != "0041":
                with gzip.GzipFile(os.path.join(HERE, f"{kind}{cn}c.d.gz"), "wb", compresslevel=9, mtime=0) as f:
                    f.write(raw)
        last = open(os.path.join(d, f"rad_detail{cn}c.plt")).read().strip().splitlines()[-1].split()
        ent["rad_detail"] = {k: (int(v) if k == "itr_gas" else float(v)) for k, v in zip(KEYS, last)}
        meta["case" + cn] = ent
    json.dump(meta, open(os.path.join(HERE, "radiation_fixtures.json"), "w"), indent=1)
    print(json.dumps(meta, indent=1))


if __name__ == "__main__":
    main(*sys.argv[1:])
