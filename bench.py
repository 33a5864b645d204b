#!/usr/bin/env python3
"""Benchmark of the SPARC2 time-step path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload sunspot_mhd] [--precision fp64]

A "step" is one pass of the driver loop body (driver.f90:123-200): one STEP
(five fused RK-stage kernels + the final update and diagnostics) and, on the
cadence steps, FILTER_VARS.  The state lives in HBM before the timed region
starts.  Prints one JSON line on stdout; progress goes to stderr.

Default workload: the sunspot MHD run (256^3, 6-face CPML, FP64) -- the
configuration the headline metric is quoted on.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gpt-updates/s & %HBM peak, FP32 vs FP64 at stated L2 err; served runs/h, p99"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def n_background(cfg):
    """3-D background fields read per stage: 0 for quiet z-profile runs, 5 for
    thermal (rho0, c2rho0, gradrho0_z, gradp0_z, sponge), 16 for MHD."""
    if cfg.mhd:
        return 16
    if cfg.magnetic == 0 and cfg.periodic and not cfg.use_hpml:
        return 0
    return 5


def psi_points(cfg):
    return sum(int(np.prod(cfg.face_shape(f))) for f in range(6) if cfg.face_active(f))


def alg_bytes_stage(cfg, w):
    """Algorithmic HBM bytes of one middle RK stage launch: read the base
    state and the previous stage RHS, write the new RHS (3 nv streams), read
    the 3-D backgrounds, and the same three streams for the CPML psi."""
    N = cfg.npoints
    return w * (N * (3 * cfg.nv + n_background(cfg)) + 3 * cfg.npsi * psi_points(cfg))


def alg_bytes_step(cfg, w):
    """B/pt x N for a whole step: 5 stages + amortised filters (+ the Kahan
    term in the FP32 form)."""
    N = cfg.npoints
    nv = cfg.nv
    per_pt = 15 * nv + 5 * n_background(cfg) + 2 * nv * (1.0 / cfg.zfreq + 1.0 / cfg.xyfreq)
    b = w * (N * per_pt + 15 * cfg.npsi * psi_points(cfg))
    if cfg.precision == "fp32":
        b += 2 * nv * 4 * N
    return b


def cpu_baseline(workload, precision, steps, threads):
    """The oracle (C restatement, OpenMP) on the same workload; only the
    time-step loop is timed."""
    from oracle.oracle import Oracle, lib
    from sparc2_amd import workloads
    cfg, bg = workloads.build(workload, precision="fp64")
    lib().orc_set_threads(threads)
    t0 = time.perf_counter()
    o = Oracle(cfg, bg)
    log(f"[cpu] oracle set-up {time.perf_counter() - t0:.1f} s")
    o.run(0, 0)                       # first step (t = 0, no filter)
    t0 = time.perf_counter()
    o.run(1, steps)
    dt = time.perf_counter() - t0
    gpts = cfg.npoints * steps / dt / 1e9
    log(f"[cpu] {steps} steps in {dt:.2f} s -> {gpts:.4f} GPt/s on {threads} threads")
    return {"value": gpts, "unit": "GPt-updates/s", "cores": threads, "kind": "port",
            "sample": f"{workload} full grid {cfg.nx}x{cfg.ny}x{cfg.nz} FP64, {steps} driver "
                      f"steps (1..{steps}) after one untimed step, set-up excluded"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="sunspot_mhd")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--profile-steps", type=int, default=4)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile", default="0,0,0")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: the whole domain lives on one device; "
              "multi-GPU sharding is not built in this round", file=sys.stderr)
        sys.exit(2)

    from sparc2_amd import workloads
    from sparc2_amd.solver import Solver

    cfg, bg = workloads.build(args.workload, precision=args.precision)
    w = 4 if args.precision == "fp32" else 8
    log(f"[bench] workload {args.workload} {cfg.nx}x{cfg.ny}x{cfg.nz} {args.precision}")
    t0 = time.perf_counter()
    tile = tuple(int(v) for v in args.tile.split(","))
    s = Solver(cfg, bg, tile=tile)
    log(f"[bench] set-up {time.perf_counter() - t0:.1f} s, tile {s.tile()}")

    W, K = args.warmup, args.steps
    s.advance(0, W - 1)
    s.sync()
    log(f"[bench] warm-up {W} steps done")
    t0 = time.perf_counter()
    s.advance(W, W + K - 1)
    s.sync()
    elapsed = time.perf_counter() - t0
    ms_step = elapsed / K * 1e3
    gpts = cfg.npoints * K / elapsed / 1e9
    d = s.diag()
    log(f"[bench] {K} steps in {elapsed:.3f} s: {ms_step:.3f} ms/step, {gpts:.3f} GPt/s, "
        f"diag {d}")

    # dominant kernel: the middle RK stages, timed with HIP events on the
    # context's stream
    P = args.profile_steps
    s.timing(True)
    s.advance(W + K, W + K + P - 1)
    s.sync()
    times = {}
    for kind, name in ((0, "stage1"), (1, "stage2_4"), (2, "stage5"), (3, "filter")):
        n, ms = s.kernel_time(kind)
        times[name] = (n, ms / n if n else None)
    s.timing(False)
    n_mid, avg_mid = times["stage2_4"]
    bytes_stage = alg_bytes_stage(cfg, w)
    achieved = bytes_stage / (avg_mid * 1e-3) / 1e9
    log(f"[bench] stage kernel avg {avg_mid:.4f} ms over {n_mid} launches, "
        f"{bytes_stage / 1e9:.3f} GB algorithmic -> {achieved:.1f} GB/s; other kernels {times}")
    step_bytes = alg_bytes_step(cfg, w)
    tile_used, nblk = s.tile()
    s.close()

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        try:
            cpu = cpu_baseline(args.workload, args.precision, args.cpu_steps, threads)
        except Exception as e:  # the GPU figure stands on its own
            log(f"[cpu] baseline failed: {e!r}")

    out = {
        "metric": METRIC,
        "value": round(gpts, 5),
        "unit": "GPt-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+kahan",
        "data": "synthetic (generated flux-tube/polytrope background, seeded-free pulse source)",
        "config": {
            "workload": args.workload,
            "grid": [cfg.nx, cfg.ny, cfg.nz],
            "physics": "MHD" if cfg.mhd else "quiet",
            "precision": args.precision,
            "dt_s": cfg.timestep,
            "pml": list(cfg.npml),
            "filters_steps": [cfg.xyfreq, cfg.zfreq],
            "step_hbm_bytes_algorithmic": int(step_bytes),
            "step_hbm_frac": round(step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "tile": list(tile_used) + [nblk],
            "kernel_avg_ms": {k: (round(v[1], 5) if v[1] else None) for k, v in times.items()},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": "k_stage (RK stages 2-4)",
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
