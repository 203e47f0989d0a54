"""Batched serving of a parameter sweep on one GPU: many furnace cases of
uneven size advanced together.

Each case replays its captured SIMPLER iteration (one HIP graph per case)
until res_mass < 1e-5 or an iteration cap.  Cases run in a pool of worker
processes that share the GPU: the solver binds a case's constants
per device in each process, so concurrent cases need separate processes;
the GPU overlaps their kernels.  Workers take cases largest first from a
shared queue (longest-processing-time order), so the pool drains evenly.
Each case first runs the launch-shape autotuner (--autotune iterations per
candidate; they count as the case's first iterations).  Reported:
aggregate cell-updates/s over the whole sweep and over the cases
that did not stall, per-case wall time to the first iteration below each
residual level of LEVELS (1e-2 ... 1e-5), the lowest level every
non-stalled case reached with each case's time to it, and the p50/p95 time
from the start of the sweep to each case's completion.  A case has stalled
when the last third of its iterations brings res_mass no lower than the
first two thirds did, or it ends above 1e-2; stalled cases are listed and
left out of the aggregate over converging cases.

With --learned the p' solve of every case starts from the network's guess
(gfm_hip_set_pguess; the 192x96x64 network for that grid, the 48x24x16 one
for the others) and is relaxed to --pcorr-tol (0.01 when not given), the
guard falling back to zero where the guess does not lower the residual.

With --pcorr-tol the p' solve of every case is relaxed to that
calc_residual tolerance from zero (GFM_PCORR_TOLERANCE), with --block-corr
each sweep preceded by the x-plane block correction; without it the solve
is the reference's fixed sweeps.

  python -m gfm_amd.serve [--cases 24 --workers 8 --cap 200 --precision mixed]
                          [--pcorr-tol 0.01 --block-corr] [--learned]
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

SIZES = [(64, 32, 24), (96, 48, 32), (128, 64, 48), (192, 96, 64), (256, 96, 64)]
THRESHOLDS = (1e-3, 1e-4, 1e-5)
LEVELS = (1e-2, 5e-3, 2e-3, 1e-3, 5e-4, 2e-4, 1e-4, 5e-5, 2e-5, 1e-5)
DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def _weights(grid):
    own = os.path.join(DATA, "pguess_furnace{}x{}x{}.npz".format(*grid))
    return own if os.path.exists(own) else os.path.join(DATA, "pguess_furnace48x24x16.npz")


def stalled(hist):
    """No progress in the last third of the iterations, or above 1e-2 at
    the end."""
    h = np.asarray(hist, dtype=float)
    if len(h) < 6:
        return False
    m = len(h) * 2 // 3
    return bool(h[-1] > 1e-2 or h[m:].min() >= h[:m].min())


def sweep(n, seed):
    """n cases: a grid size drawn from SIZES, burner velocity U[20,40] m/s,
    air preheat U[900,1300] K."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(n):
        nx, ny, nz = SIZES[rng.integers(len(SIZES))]
        out.append(dict(index=q, grid=(int(nx), int(ny), int(nz)), fuel_u=float(rng.uniform(20.0, 40.0)),
                        air_t=float(rng.uniform(900.0, 1300.0)), seed=int(1000 + q)))
    return out


def _run_case(job, cfg, t_sweep0):
    from . import case as cs
    from .solver import CombustionSpace
    nx, ny, nz = job["grid"]
    c = cs.furnace_large(nx=nx, ny=ny, nz=nz, seed=job["seed"], maxgi=10 ** 7, bgcon=min(THRESHOLDS),
                         fuel_u=job["fuel_u"], air_t=job["air_t"])
    k = c.ctrl
    k.sweep_order = 1
    k.want_pre_residuals = 0
    k.precision = 1 if cfg["precision"] == "mixed" else 0
    if cfg.get("pcorr_tol", 0.0) > 0.0:
        k.pcorr_mode, k.pcorr_tol, k.pcorr_max_sweeps = 1, cfg["pcorr_tol"], 400
        k.pcorr_block_corr = 1 if cfg.get("block_corr") else 0
    net = None
    if cfg.get("learned"):
        from . import pguess
        k.pcorr_mode, k.pcorr_tol, k.pcorr_max_sweeps = 1, cfg.get("pcorr_tol") or 0.01, 400
        k.pcorr_guess = 1
        k.pcorr_guess_accept = 0.5
        net = pguess.load(_weights(job["grid"]), max_cells=nx * ny * nz)
    first = {f"{t:g}": None for t in LEVELS}
    t0 = time.perf_counter()
    n = 0
    res = float("nan")
    hist = []
    fallbacks = 0
    try:
        with CombustionSpace(c) as s:
            if net is not None:
                s.set_pguess(net)
            if cfg.get("autotune", 0) > 0 and cfg["cap"] > 6 * cfg["autotune"]:
                # the autotuner's iterations are ordinary iterations of the case
                s.autotune(cfg["autotune"])
                n = 6 * cfg["autotune"]
                res = float(k.res_mass)
                hist.append(res)
            while n < cfg["cap"]:
                h = s.sphase(min(10, cfg["cap"] - n))
                if len(h) == 0:
                    break
                fallbacks += int(k.pcorr_fallback)
                for r in h:
                    n += 1
                    res = float(r)
                    hist.append(res)
                    for t in LEVELS:
                        if first[f"{t:g}"] is None and res < t:
                            first[f"{t:g}"] = {"iterations": n, "seconds": round(time.perf_counter() - t0, 3)}
                if res < min(THRESHOLDS):
                    break
    finally:
        if net is not None:
            net.close()
    t1 = time.perf_counter()
    return dict(index=job["index"], grid=list(job["grid"]), cells=nx * ny * nz, iterations=n, seconds=round(t1 - t0, 3),
                res_mass=res, min_res_mass=float(min(hist)) if hist else float("nan"), stalled=stalled(hist),
                first_below=first, done_at=round(time.time() - t_sweep0, 3), worker=os.getpid(),
                guess_fallbacks_sampled=fallbacks if net is not None else None)


def _worker(jobs, results, cfg, t_sweep0):
    while True:
        job = jobs.get()
        if job is None:
            break
        try:
            results.put(_run_case(job, cfg, t_sweep0))
        except Exception as e:  # report and keep serving
            results.put(dict(index=job["index"], error=f"{type(e).__name__}: {e}"))


def lowest_common_level(recs):
    """The lowest residual level of LEVELS that every record reached, with
    each record's (iterations, seconds) to it; (None, {}) if none."""
    best = None
    for t in LEVELS:
        if recs and all(r["first_below"][f"{t:g}"] for r in recs):
            best = t
    if best is None:
        return None, {}
    return best, {r["index"]: r["first_below"][f"{best:g}"] for r in recs}


def serve(cases, workers, cap, precision, log=print, pcorr_tol=0.0, block_corr=False, learned=False, autotune=0):
    """Run the sweep; returns the summary and the per-case records."""
    ctx = mp.get_context("spawn")
    jobs, results = ctx.Queue(), ctx.Queue()
    for job in sorted(cases, key=lambda j: -np.prod(j["grid"])):
        jobs.put(job)
    for _ in range(workers):
        jobs.put(None)
    cfg = dict(cap=cap, precision=precision, pcorr_tol=pcorr_tol, block_corr=block_corr, learned=learned,
               autotune=autotune)
    t0w = time.time()
    t0 = time.perf_counter()
    procs = [ctx.Process(target=_worker, args=(jobs, results, cfg, t0w)) for _ in range(workers)]
    for p in procs:
        p.start()
    recs = []
    while len(recs) < len(cases):
        r = results.get()
        recs.append(r)
        if "error" in r:
            log(f"case {r['index']} failed: {r['error']}")
        else:
            log(f"case {r['index']} {r['grid']}: {r['iterations']} iterations in {r['seconds']:.1f} s, "
                f"res_mass {r['res_mass']:.2e} (lowest {r['min_res_mass']:.2e}){' STALLED' if r['stalled'] else ''}, "
                f"done at {r['done_at']:.1f} s")
    wall = time.perf_counter() - t0
    for p in procs:
        p.join()
    ok = [r for r in recs if "error" not in r]
    moving = [r for r in ok if not r["stalled"]]
    work = sum(r["cells"] * r["iterations"] for r in ok)
    work_moving = sum(r["cells"] * r["iterations"] for r in moving)
    done = np.array([r["done_at"] for r in ok]) if ok else np.array([np.nan])
    level, times = lowest_common_level(moving)
    summary = dict(cases=len(cases), failed=len(recs) - len(ok), workers=workers, precision=precision, cap=cap,
                   pcorr_tol=(pcorr_tol or (0.01 if learned else 0.0)) or None,
                   block_correction=bool(block_corr and pcorr_tol), learned_guess=bool(learned),
                   autotune_iterations_per_candidate=autotune,
                   wall_seconds=round(wall, 3), cell_updates=work,
                   aggregate_mcell_updates_per_s=round(work / wall / 1e6, 3),
                   stalled_cases=[r["index"] for r in ok if r["stalled"]],
                   aggregate_mcell_updates_per_s_converging=round(work_moving / wall / 1e6, 3),
                   completion_p50_s=round(float(np.percentile(done, 50)), 3),
                   completion_p95_s=round(float(np.percentile(done, 95)), 3),
                   lowest_common_residual=level,
                   time_to_lowest_common_residual={str(q): v for q, v in sorted(times.items())},
                   reached={f"{t:g}": sum(1 for r in ok if r["first_below"][f"{t:g}"]) for t in THRESHOLDS})
    return summary, sorted(recs, key=lambda r: r["index"])


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m gfm_amd.serve")
    ap.add_argument("--cases", type=int, default=24)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--cap", type=int, default=200, help="iteration cap per case")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--pcorr-tol", type=float, default=0.0,
                    help="relax p' to this calc_residual from zero (0: the reference's fixed sweeps)")
    ap.add_argument("--block-corr", action="store_true", help="x-plane block correction before each p' sweep")
    ap.add_argument("--learned", action="store_true", help="the learned p' guess (p' relaxed to --pcorr-tol, 0.01)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="iterations per candidate of each case's launch-shape autotuner (0: default shapes)")
    a = ap.parse_args(argv)
    if a.workers < 1 or a.workers > 15:
        print("serve: 1..15 worker processes share the GPU", file=sys.stderr)
        return 2
    if a.block_corr and not a.pcorr_tol > 0.0:
        print("serve: --block-corr needs --pcorr-tol", file=sys.stderr)
        return 2
    summary, recs = serve(sweep(a.cases, a.seed), a.workers, a.cap, a.precision,
                          log=lambda m: print(m, flush=True), pcorr_tol=a.pcorr_tol, block_corr=a.block_corr,
                          learned=a.learned, autotune=a.autotune)
    print(json.dumps({"summary": summary, "cases": recs}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
