"""Learned initial guess of the pressure-correction (p') solve: training
pairs dumped by the solver, training on the GPU, and evaluation against the
zero guess at the same p' tolerance.

The reference starts every p' solve from zero and relaxes 5 sweeps
(C/Gslv3v.f90:149-164).  In GFM_PCORR_TOLERANCE mode the solver relaxes
to calc_residual <= pcorr_tol instead, from zero or from this network's
prediction (include/gfm_hip.h, include/gfm_pnet.h).  Pairs are (BS/AP and
the open-row mask, p' after a tight solve) from the furnace at seeded
operating points; each sample is normalised by the root mean square over
open rows (input and target separately: the solver rescales the prediction
by the least-squares factor alpha, so only its shape matters).

Command line (one GPU):
  python -m gfm_amd.pguess train --out W.npz [--points 32 --pairs 128 --epochs 20]
  python -m gfm_amd.pguess eval --weights W.npz [--points 8 --iters 60]
"""
import argparse
import json
import sys
import time

import numpy as np

from . import case as cs
from .pnet import DEFAULT_CHANNELS, PNet
from .solver import CombustionSpace


def operating_points(n, seed):
    """Burner velocity U[20,40] m/s, air preheat U[900,1300] K, excess air
    U[5,20] %."""
    rng = np.random.default_rng(seed)
    return [dict(fuel_u=float(rng.uniform(20.0, 40.0)), air_t=float(rng.uniform(900.0, 1300.0)),
                 excess_air=float(rng.uniform(0.05, 0.20))) for _ in range(n)]


def normalise(x, p=None):
    """Scale x[...,0] (and p) to unit root mean square over open rows."""
    x = x.copy()
    m = x[..., 1] > 0
    rx = float(np.sqrt(np.mean(x[..., 0][m].astype(np.float64) ** 2))) if m.any() else 1.0
    x[..., 0] /= rx if rx > 0 else 1.0
    if p is None:
        return x
    rp = float(np.sqrt(np.mean(p[m].astype(np.float64) ** 2))) if m.any() else 1.0
    return x, p / (rp if rp > 0 else 1.0)


def make_case(op, grid=(48, 24, 16), seed=100, large=False, precision=0, tol=1e-3, max_sweeps=400, guess=0, dump=0):
    nx, ny, nz = grid
    fn = cs.furnace_large if large else cs.furnace_small
    c = fn(nx=nx, ny=ny, nz=nz, seed=seed, maxgi=10 ** 7, bgcon=0.0, **op)
    k = c.ctrl
    k.sweep_order = 1
    k.want_pre_residuals = 0
    k.precision = precision
    k.pcorr_mode = 1
    k.pcorr_guess = guess
    k.pcorr_tol = tol
    k.pcorr_max_sweeps = max_sweeps
    k.pcorr_dump = dump
    return c


def generate_pairs(points, grid=(48, 24, 16), iters=130, skip=2, tol=1e-3, max_sweeps=400, seed0=100, log=None,
                   large=False, every=1):
    """Run each operating point `iters` SIMPLER iterations with the p' solve
    relaxed to `tol` from zero; keep the pair of every `every`-th iteration
    after the first `skip`.  Returns X [N,nz,ny,nx,2], T [N,nz,ny,nx]
    (normalised)."""
    X, T = [], []
    for q, op in enumerate(points):
        c = make_case(op, grid, seed=seed0 + q, large=large, tol=tol, max_sweeps=max_sweeps, dump=1)
        t0 = time.perf_counter()
        sw = []
        with CombustionSpace(c) as s:
            for it in range(iters):
                s.sphase(1)
                sw.append(c.ctrl.pcorr_sweeps)
                if it >= skip and (it - skip) % every == 0:
                    x, p = s.pcorr_pair()
                    xn, pn = normalise(x, p)
                    X.append(xn)
                    T.append(pn)
        if log:
            log(f"point {q} {op}: {iters} iterations in {time.perf_counter() - t0:.1f} s, p' sweeps mean "
                f"{np.mean(sw):.1f}, res_mass {c.ctrl.res_mass:.3e}")
    return np.stack(X).astype(np.float32), np.stack(T).astype(np.float32)


def save(net, path, meta=None):
    np.savez(path, params=net.get_params(), channels=np.array(net.channels, dtype=np.int32),
             meta=json.dumps(meta or {}))


def load(path, max_cells, max_batch=1, device=0):
    with np.load(path, allow_pickle=False) as f:
        params, ch = f["params"], tuple(int(c) for c in f["channels"])
    net = PNet(ch, max_batch=max_batch, max_cells=max_cells, device=device)
    net.set_params(params)
    return net


def run_point(op, net, guess, grid, iters, tol, max_sweeps, large=False, precision=0, seed=200):
    """One operating point: `iters` iterations with the p' solve relaxed to
    tol from zero (guess 0) or from the network (guess 1).  Returns the
    per-iteration p' sweeps, fallbacks, res_mass and the wall time."""
    c = make_case(op, grid, seed=seed, large=large, precision=precision, tol=tol, max_sweeps=max_sweeps, guess=guess)
    sw, fb, rm = [], [], []
    with CombustionSpace(c) as s:
        if guess:
            s.set_pguess(net)
        s.sphase(1)  # capture the graphs outside the timed region
        sw.append(c.ctrl.pcorr_sweeps); fb.append(c.ctrl.pcorr_fallback); rm.append(c.ctrl.res_mass)
        t0 = time.perf_counter()
        for _ in range(iters - 1):
            s.sphase(1)
            sw.append(c.ctrl.pcorr_sweeps); fb.append(c.ctrl.pcorr_fallback); rm.append(c.ctrl.res_mass)
        dt = time.perf_counter() - t0
    return dict(sweeps=sw, fallback=fb, res_mass=rm, seconds=dt)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m gfm_amd.pguess")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", required=True)
    t.add_argument("--points", type=int, default=32)
    t.add_argument("--pairs", type=int, default=128)
    t.add_argument("--epochs", type=int, default=20)
    t.add_argument("--batch", type=int, default=16)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--tol", type=float, default=1e-3)
    t.add_argument("--seed", type=int, default=1)
    t.add_argument("--large", action="store_true", help="pairs from the stretched 4-burner furnace instead of the "
                                                        "single-burner chamber")
    t.add_argument("--every", type=int, default=1, help="keep every n-th iteration's pair (pairs per point unchanged)")
    t.add_argument("--grid", type=int, nargs=3, default=None, help="nx ny nz (default 48 24 16, or 192 96 64 with --large)")
    e = sub.add_parser("eval")
    e.add_argument("--weights", required=True)
    e.add_argument("--points", type=int, default=8)
    e.add_argument("--iters", type=int, default=60)
    e.add_argument("--tol", type=float, default=1e-3)
    e.add_argument("--large", action="store_true", help="192x96x64 stretched furnace instead of 48x24x16")
    e.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    if a.cmd == "train":
        pts = operating_points(a.points, a.seed)
        t0 = time.perf_counter()
        grid = tuple(a.grid) if a.grid else ((192, 96, 64) if a.large else (48, 24, 16))
        X, T = generate_pairs(pts, grid=grid, iters=a.pairs * a.every + 2, skip=2, tol=a.tol, log=log, large=a.large,
                              every=a.every)
        log(f"{len(X)} pairs in {time.perf_counter() - t0:.1f} s")
        net = PNet(DEFAULT_CHANNELS, max_batch=a.batch, max_cells=int(np.prod(X.shape[1:4])), seed=a.seed)
        log(f"network: {net.n_effective} parameters ({net.n_params} with padding)")
        t0 = time.perf_counter()
        hist = []
        for ep in range(a.epochs):  # one call per epoch, so progress is visible
            hist.append(float(net.train(X, T, epochs=1, batch=a.batch, lr=a.lr, seed=a.seed * 1000 + ep)[0]))
            log(f"epoch {ep}: loss {hist[-1]:.4g} ({time.perf_counter() - t0:.1f} s)")
        hist = np.array(hist)
        dt = time.perf_counter() - t0
        log(f"trained {a.epochs} epochs in {dt:.1f} s, loss per epoch {np.array2string(hist, precision=4)}")
        save(net, a.out, dict(points=a.points, pairs=len(X), epochs=a.epochs, batch=a.batch, lr=a.lr, tol=a.tol,
                              seed=a.seed, loss=hist.tolist(), train_seconds=dt, grid=list(grid),
                              geometry="large" if a.large else "small", every=a.every))
        net.close()
    else:
        grid = (192, 96, 64) if a.large else (48, 24, 16)
        pts = operating_points(a.points, 1000 + a.points)  # held out: a different seed stream
        net = load(a.weights, max_cells=int(np.prod(grid)))
        prec = 1 if a.precision == "mixed" else 0
        rows = []
        for q, op in enumerate(pts):
            z = run_point(op, net, 0, grid, a.iters, a.tol, 1000, a.large, prec, seed=500 + q)
            l = run_point(op, net, 1, grid, a.iters, a.tol, 1000, a.large, prec, seed=500 + q)
            rows.append(dict(op=op, zero=z, learned=l))
            log(f"point {q}: p' sweeps zero {np.mean(z['sweeps']):.1f} learned {np.mean(l['sweeps']):.1f} "
                f"(fallback {int(np.sum(l['fallback']))}/{a.iters}); time zero {z['seconds']:.2f} s learned "
                f"{l['seconds']:.2f} s; res_mass {z['res_mass'][-1]:.3e} / {l['res_mass'][-1]:.3e}")
        sz = np.mean([np.mean(r["zero"]["sweeps"]) for r in rows])
        sl = np.mean([np.mean(r["learned"]["sweeps"]) for r in rows])
        tz = sum(r["zero"]["seconds"] for r in rows)
        tl = sum(r["learned"]["seconds"] for r in rows)
        log(json.dumps(dict(grid=grid, precision=a.precision, tol=a.tol, iters=a.iters, points=a.points,
                            mean_sweeps_zero=sz, mean_sweeps_learned=sl, seconds_zero=tz, seconds_learned=tl)))
        net.close()


if __name__ == "__main__":
    sys.exit(main())
