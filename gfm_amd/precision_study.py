"""Error of the FP32 sweeps against the FP64 GPU run (mixed precision: FP32
assembly and line sweeps; p' defect, balances and h->T in FP64).

The furnace is developed in FP64 for `dev` iterations, then three runs
continue from that one state for `iters` iterations: FP64, FP32 sweeps, and
FP64 with T perturbed by 1e-7 (relative, seeded).  The third run measures
how strongly the flow itself amplifies a rounding-size perturbation over
the same iterations, the floor below which no precision comparison can go.
Reported per quantity: max and RMS relative error against the FP64 run
(velocities on their open faces, scalars on open cells; the scale is the
field's max |value| for the max error and its RMS for the RMS error) and the
relative error of the convective heat to the crown (top faces), the floor
or glass (bottom faces) and all walls.

  python -m gfm_amd.precision_study --grid 384 160 96 --dev 300 --iters 50
"""
import argparse
import json
import sys
import time

import numpy as np

from . import case as cs
from .solver import CombustionSpace

SCALARS = {"T": ("t", None), "p": ("p", None), "Y_fuel": ("gf", cs.IYF), "Y_O2": ("gf", cs.IYO2),
           "Y_CO2": ("gf", cs.IYCO2), "Y_H2O": ("gf", cs.IYH2O), "Y_N2": ("gf", cs.IYN2)}


def _open_centres(ib):
    m = np.zeros(ib.shape, dtype=bool)
    m[1::2, 1::2, 1::2] = ib[1::2, 1::2, 1::2] <= 0
    return m


def _open_faces(ib, iu):
    m = np.zeros(ib.shape, dtype=bool)
    sl = [slice(1, None, 2)] * 3
    sl[iu - 1] = slice(2, None, 2)
    sl = tuple(sl)
    m[sl] = ib[sl] <= 0
    return m


def _err(ref, x):
    d = x - ref
    mx = float(np.max(np.abs(d)) / max(np.max(np.abs(ref)), 1e-300))
    rms = float(np.sqrt(np.mean(d * d)) / max(np.sqrt(np.mean(ref * ref)), 1e-300))
    return {"max": mx, "rms": rms}


def compare(ref, other):
    ib = ref.arrays["ibcell"]
    mc = _open_centres(ib)
    out = {}
    for name, (arr, l) in SCALARS.items():
        a = ref.arrays[arr] if l is None else ref.arrays[arr][..., l - 1]
        b = other.arrays[arr] if l is None else other.arrays[arr][..., l - 1]
        out[name] = _err(a[mc], b[mc])
    for iu in (1, 2, 3):
        mf = _open_faces(ib, iu)
        out["uvw"[iu - 1]] = _err(ref.arrays["ug"][..., iu - 1][mf], other.arrays["ug"][..., iu - 1][mf])
    for q in ("q_wall_top", "q_wall_bottom", "q_wall", "res_mass"):
        a, b = getattr(ref.ctrl, q), getattr(other.ctrl, q)
        out[q] = {"ref": a, "rel": abs(b - a) / max(abs(a), 1e-300)}
    return out


def run(grid, dev, iters, seed, log=print):
    rng = np.random.default_rng(seed)
    op = dict(fuel_u=float(rng.uniform(20.0, 40.0)), air_t=float(rng.uniform(900.0, 1300.0)))
    nx, ny, nz = grid
    t0 = time.perf_counter()
    base = cs.furnace_small(nx=nx, ny=ny, nz=nz, seed=seed, maxgi=10 ** 7, bgcon=0.0, **op)
    base.ctrl.sweep_order = 1
    base.ctrl.want_pre_residuals = 0
    log(f"case {nx}x{ny}x{nz} ({nx * ny * nz} cells) {op} built in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    with CombustionSpace(base) as s:
        for q in range(0, dev, 50):
            h = s.sphase(min(50, dev - q))
            log(f"  developing FP64: iteration {base.ctrl.itr_gas}, res_mass {h[-1]:.3e}")
        s.download()
    log(f"developed in {time.perf_counter() - t0:.1f} s")
    runs = {}
    for name in ("fp64", "fp32", "fp64_perturbed"):
        c = base.clone()
        c.ctrl.precision = 1 if name == "fp32" else 0
        if name == "fp64_perturbed":
            p = np.random.default_rng(seed + 1).standard_normal(c.arrays["t"].shape)
            c.arrays["t"][...] *= 1.0 + 1e-7 * p
        t0 = time.perf_counter()
        with CombustionSpace(c, initsv=False) as s:
            h = s.sphase(iters)
            s.download()
        log(f"  {name}: {len(h)} iterations in {time.perf_counter() - t0:.1f} s, res_mass {h[-1]:.3e}")
        runs[name] = c
        if name != "fp64":
            del c
    ref = runs.pop("fp64")
    result = {"grid": list(grid), "cells": nx * ny * nz, "operating_point": op, "developed_iterations": dev,
              "compared_iterations": iters}
    for name, c in runs.items():
        result[name] = compare(ref, c)
    return result


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m gfm_amd.precision_study")
    ap.add_argument("--grid", type=int, nargs=3, default=(384, 160, 96))
    ap.add_argument("--dev", type=int, default=300)
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--seed", type=int, default=3)
    a = ap.parse_args(argv)
    res = run(tuple(a.grid), a.dev, a.iters, a.seed, log=lambda m: print(m, flush=True))
    print(json.dumps(res))


if __name__ == "__main__":
    sys.exit(main())
