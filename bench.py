#!/usr/bin/env python3
"""Benchmark of the combustion-space SIMPLER path on one MI355X.

A step is one SIMPLER outer iteration (one GFLOW: momentum, pressure, p',
velocity correction, 7 scalar equations, properties, boundary updates)
over the 192x96x64 stretched furnace (1,179,648 cells), all state resident
in HBM.  Prints one JSON line with Mcell-updates/s (cells x iterations /
s), the roofline figure of the line-relaxation kernel measured with HIP
events during the timed region, and the CPU baseline (the oracle, one core,
on a bounded sample of the same workload).  Beside the plain FP64 value it
reports FP32 configurations with the p' solve relaxed to a tolerance (from
zero, from the learned guess, with the x-plane block correction) and the
time each takes to reach res_mass 1e-3, 1e-4, 5e-5 and 1e-5.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic byte model of one cell through one SIMPLER outer iteration
# (words of 8 bytes in FP64, 4 in FP32 for the assembly and sweeps): 198
# words of assembly, 820 of line relaxation (41 sweeps x 2 line directions
# x (8 coefficients + 1 read + 1 write)), 108 of post-solve residuals, 12 of
# velocity correction, 8 of interpolation, plus 88 bytes of FP64 property
# update.  FP64: 9,256 B per cell-update; mixed: about 4,692 B.
MODEL_BYTES = {"fp64": 1146 * 8 + 88, "mixed": 4692}
LINE_WORDS = 820          # line-relaxation term per cell-iteration
LINE_LAUNCHES_PER_SWEEP = 8  # zebra order: 4 colour visits x 2 passes, each launch one colour (half the cells)
PMC_SUMMARY = os.path.join("profiles", "pmc_k_lines_furnace192x96x64.md")


def measured_traffic(precision, dims):
    """HBM bytes per k_lines launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (counters cannot be read from inside this process).
    Only for the workload the counters were taken on."""
    if dims != (192, 96, 64):
        return None
    try:
        with open(os.path.join(ROOT, PMC_SUMMARY)) as f:
            for ln in f:
                if ln.startswith(f"traffic_bytes_per_launch[{precision}] = "):
                    return float(ln.split("=")[1])
    except OSError:
        pass
    return None


def cpu_baseline(spec_args, iters=2):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from gfm_amd import case as cs
    c = cs.furnace_large(**spec_args)
    c.ctrl.sweep_order = 0
    c.ctrl.want_pre_residuals = 0
    oracle.initsv_tail(c)
    t = time.perf_counter()
    n = len(oracle.sphase(c, iters))
    dt = time.perf_counter() - t
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": c.ncells * n / dt / 1e6, "unit": "Mcell-updates/s", "cores": 1, "kind": "port",
            "cpu_model": model, "nproc": os.cpu_count(),
            "sample": f"{n} SIMPLER iterations of the same {c.ctrl.mp // 2 - 2}x{c.ctrl.np // 2 - 2}x"
                      f"{c.ctrl.lp // 2 - 2} case, reference x-plane order, "
                      f"oracle/gfm_oracle.c (gcc -O2) on 1 host core, {dt:.1f} s"}


PGUESS_DIR = os.path.join(ROOT, "gfm_amd", "data")
THRESHOLDS = (1e-3, 1e-4, 5e-5, 1e-5)


def configure(case, mode, pcorr_tol):
    """fp64: the reference algorithm (FP64, p' from zero, 5 sweeps).
    mixed: FP32 assembly and sweeps, the reference's p' solve.
    mixed_zero: FP32, p' from zero relaxed to pcorr_tol.
    learned: FP32, p' from the network's guess relaxed to pcorr_tol.
    A suffix _bc adds the x-plane block correction before each p' sweep
    (mixed_zero_bc, learned_bc; fp64_zero_bc is FP64 with p' from zero
    relaxed to pcorr_tol with it)."""
    k = case.ctrl
    k.sweep_order = 1
    k.want_pre_residuals = 0
    base, bc = (mode[:-3], 1) if mode.endswith("_bc") else (mode, 0)
    if base == "fp64":
        k.precision, k.pcorr_mode = 0, 0
    elif base == "mixed":
        k.precision, k.pcorr_mode = 1, 0
    else:
        k.precision = 0 if base == "fp64_zero" else 1
        k.pcorr_mode, k.pcorr_guess = 1, 1 if base == "learned" else 0
        k.pcorr_tol, k.pcorr_max_sweeps = pcorr_tol, 400
        k.pcorr_guess_accept = 0.5  # keep a guess only if it halves the start residual
        k.pcorr_block_corr = bc


def throughput(spec_args, mode, net, steps, warmup, pcorr_tol, sync, autotune=0):
    """W untimed iterations, then K timed ones (the library synchronises its
    stream at the end of every iteration, when it reads the residuals).
    With autotune > 0 the launch shapes of the per-position box kernels are
    chosen first (6*autotune untimed iterations); returns their table."""
    from gfm_amd import case as cs
    from gfm_amd.solver import CombustionSpace
    case = cs.furnace_large(**spec_args)
    configure(case, mode, pcorr_tol)
    cspace = CombustionSpace(case, device=0)
    if net is not None:
        cspace.set_pguess(net)
    cspace.tuning = cspace.autotune(autotune) if autotune > 0 else ""
    for _ in range(warmup):
        cspace.gflow()
    sync()
    t0 = time.perf_counter()
    hist = cspace.sphase(steps)
    sync()
    dt = time.perf_counter() - t0
    return cspace, case, hist, dt


def time_to_residual(spec_args, mode, net, cap, pcorr_tol):
    """Iterate from the initial state until res_mass < 1e-5 or `cap`
    iterations; wall time and iteration count at the first crossing of each
    threshold (the graph capture of the first iteration included)."""
    from gfm_amd import case as cs
    from gfm_amd.solver import CombustionSpace
    case = cs.furnace_large(**dict(spec_args, maxgi=cap, bgcon=min(THRESHOLDS)))
    configure(case, mode, pcorr_tol)
    out = {f"{t:g}": None for t in THRESHOLDS}
    best = float("inf")
    with CombustionSpace(case, device=0) as s:
        if net is not None:
            s.set_pguess(net)
        t0 = time.perf_counter()
        n = sweeps = 0
        while n < cap:
            h = s.sphase(1)
            if len(h) == 0:
                break
            n += 1
            r = float(h[0])
            best = min(best, r)
            sweeps += case.ctrl.pcorr_sweeps
            for t in THRESHOLDS:
                if out[f"{t:g}"] is None and r < t:
                    out[f"{t:g}"] = {"seconds": round(time.perf_counter() - t0, 3), "iterations": n}
            if r < min(THRESHOLDS):
                break
        total = time.perf_counter() - t0
    return {"first_below": out, "iterations_run": n, "seconds": round(total, 3), "min_res_mass": best,
            "mean_pcorr_sweeps": round(sweeps / max(n, 1), 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=192)
    ap.add_argument("--ny", type=int, default=96)
    ap.add_argument("--nz", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=2)
    ap.add_argument("--eager", action="store_true", help="launch kernels one by one instead of replaying a graph")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned+FP32 configuration")
    ap.add_argument("--pcorr-tol", type=float, default=0.01,
                    help="p' tolerance of the learned+FP32 configuration (calc_residual)")
    ap.add_argument("--pcorr-tol-bc", type=float, default=0.05,
                    help="p' tolerance of FP32 from zero with the block correction (at 0.02-0.1 every solve is one "
                         "corrected sweep)")
    ap.add_argument("--pcorr-tol-tight", type=float, default=1e-3,
                    help="p' tolerance of the second set of time-to-residual runs (FP32 from zero, learned+FP32, "
                         "learned+FP32 with the block correction); 0 skips them")
    ap.add_argument("--ttr-cap", type=int, default=800,
                    help="iteration cap of the time-to-residual runs (0 skips them)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="iterations per candidate of the launch-shape autotuner before the warm-up (0: default shapes)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run a torch device copy of this size (reference kernel for rocprof byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "a case shards as x-slabs across GPUs only in the design notes", file=sys.stderr)
        sys.exit(2)

    try:
        import torch
        sync = torch.cuda.synchronize
    except Exception:  # the library synchronises its own stream per iteration
        sync = lambda: None

    if args.eager:
        os.environ["GFM_HIP_GRAPH"] = "0"
    if args.copy_calibration_mib:
        import torch
        src = torch.ones(args.copy_calibration_mib * 131072, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
    spec_args = dict(nx=args.nx, ny=args.ny, nz=args.nz, seed=2, maxgi=10 ** 7, bgcon=0.0)
    n = args.nx * args.ny * args.nz
    net = None
    weights = os.path.join(PGUESS_DIR, f"pguess_furnace{args.nx}x{args.ny}x{args.nz}.npz")
    if not os.path.exists(weights):
        weights = os.path.join(PGUESS_DIR, "pguess_furnace48x24x16.npz")
    if not args.no_learned and os.path.exists(weights) and not (args.nx | args.ny | args.nz) & 7:
        from gfm_amd import pguess
        net = pguess.load(weights, max_cells=n)
    # the workload of the metric's value: plain FP64 (or FP32 sweeps with
    # the reference's p' solve, --precision mixed)
    cspace, case, hist, dt = throughput(spec_args, args.precision, None, args.steps, args.warmup, args.pcorr_tol, sync,
                                        args.autotune)
    shapes = cspace.tuning
    # per-kernel HIP-event timing (kernels launched one by one on the
    # library's stream) over a short pass right after the timed region
    cspace.kernel_times(1)
    cspace.sphase(max(2, min(5, args.steps)))
    kt = cspace.kernel_times(0)
    cspace.close()
    steps = len(hist)
    res_final = float(hist[-1]) if steps else float("nan")
    value = n * steps / dt / 1e6

    ms, by, la = kt.get("k_lines", (0.0, 0.0, 0))
    avg_s = ms / 1e3 / max(la, 1)
    word = 8 if args.precision == "fp64" else 4
    # the line term of the byte model per launch: 20 words per cell and
    # sweep (2 directions x 10 words), spread over the 8 launches of a sweep
    # of which each covers half the cells (one colour)
    per_launch = LINE_WORDS / 41 / LINE_LAUNCHES_PER_SWEEP * 2 * (n / 2) * word
    achieved = per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    kernel_model = by / max(la, 1)  # the kernel's own count: coefficients re-read by every pass
    traffic = measured_traffic(args.precision, (args.nx, args.ny, args.nz))
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": "k_lines", "avg_launch_us": round(avg_s * 1e6, 2), "bytes_per_launch": round(per_launch),
            "launches": la, "traffic_source": PMC_SUMMARY if traffic else None,
            "kernel_model_bytes_per_launch": round(kernel_model),
            "kernel_model_gbs": round(kernel_model / avg_s / 1e9, 1) if avg_s > 0 else None}
    line = {
        "metric": "time to residual 1e-5 & Mcell-updates/s: learned+FP32 vs plain FP64, 1 GPU",
        "value": round(value, 3), "unit": "Mcell-updates/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": round(dt / max(steps, 1) * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32+f64", "data": "synthetic",
        "config": {"workload": f"{'plain FP64' if args.precision == 'fp64' else 'mixed FP32/FP64'} SIMPLER, "
                               f"stretched furnace {args.nx}x{args.ny}x{args.nz} "
                               f"({n} cells), zebra x-plane order", "cells": n, "res_mass_last": res_final},
        "roofline": roof,
        "autotune": {"iterations_per_candidate": args.autotune,
                     "launch_shapes": [ln for ln in shapes.splitlines()[1:]] if shapes else None},
        # the whole iteration against the byte model (9,256 B per cell-update in FP64)
        "iteration_roofline": {"model_bytes_per_cell_update": MODEL_BYTES[args.precision],
                               "achieved": round(MODEL_BYTES[args.precision] * n * steps / dt / 1e9, 1),
                               "unit": "GB/s", "peak": HBM_PEAK_GBS,
                               "frac": round(MODEL_BYTES[args.precision] * n * steps / dt / 1e9 / HBM_PEAK_GBS, 4),
                               "ceiling_mcell_updates_s": round(HBM_PEAK_GBS * 1e9 / MODEL_BYTES[args.precision] / 1e6, 1)},
    }
    others = [("fp32", "mixed", None)]
    if net is not None:
        others += [("fp32_zero_guess_tol", "mixed_zero", None), ("learned_fp32", "learned", net)]
    others += [("fp32_zero_guess_tol_bc", "mixed_zero_bc", None)]
    def tol_of(mode):
        return args.pcorr_tol_bc if mode == "mixed_zero_bc" else args.pcorr_tol

    for name, mode, nn in others:
        c2, case2, h2, dt2 = throughput(spec_args, mode, nn, args.steps, args.warmup, tol_of(mode), sync, args.autotune)
        c2.close()
        line[name] = {"value": round(n * len(h2) / dt2 / 1e6, 3), "unit": "Mcell-updates/s",
                      "ms_per_step": round(dt2 / max(len(h2), 1) * 1e3, 3)}
        if mode != "mixed":
            line[name].update(pcorr_tol=tol_of(mode), pcorr_sweeps_last=case2.ctrl.pcorr_sweeps)
        if mode.endswith("_bc"):
            line[name]["block_correction"] = True
    if net is not None:
        line["learned_fp32"]["network"] = os.path.relpath(weights, ROOT)
    if args.ttr_cap > 0:
        ttr = {"cap_iterations": args.ttr_cap, "pcorr_tol": args.pcorr_tol}
        for name, mode, nn in [("plain_fp64", "fp64", None)] + others:
            ttr[name] = time_to_residual(spec_args, mode, nn, args.ttr_cap, tol_of(mode))
            ttr[name]["block_correction"] = mode.endswith("_bc")
            ttr[name]["pcorr_tol"] = tol_of(mode) if mode not in ("fp64", "mixed") else None
        if net is not None and args.pcorr_tol_tight > 0:
            # the p' solve relaxed further: ~100-150 sweeps per solve, where
            # the guess saves more than its inference costs
            ttr["pcorr_tol_tight"] = args.pcorr_tol_tight
            for name, mode in (("fp32_zero_guess_tol_tight", "mixed_zero"), ("learned_fp32_tol_tight", "learned"),
                               ("learned_fp32_tol_tight_bc", "learned_bc")):
                ttr[name] = time_to_residual(spec_args, mode, net if mode.startswith("learned") else None,
                                             args.ttr_cap, args.pcorr_tol_tight)
                ttr[name]["block_correction"] = mode.endswith("_bc")
        line["time_to_residual"] = ttr
        # a configuration whose run from the initial state never passes 1e-3 has stalled: its throughput
        # counts iterations that do not converge
        for name, _, _ in others:
            line[name]["reaches_1e-3"] = ttr[name]["first_below"]["0.001"] is not None
    if net is not None:
        net.close()
    kt_sum = {k: {"ms": round(v[0], 3), "launches": v[2]} for k, v in sorted(kt.items(), key=lambda kv: -kv[1][0])}
    print(json.dumps({"kernel_times_ms": kt_sum}), file=sys.stderr)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(spec_args, args.cpu_iters)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
