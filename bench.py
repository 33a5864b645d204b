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


def psi_points(cfg, faces=range(6)):
    return sum(int(np.prod(cfg.face_shape(f))) for f in faces if cfg.face_active(f))


def alg_bytes_stage(cfg, w):
    """Algorithmic HBM bytes of one middle RK stage launch: read the base
    state and the previous stage RHS, write the new RHS (3 nv streams), read
    the 3-D backgrounds, and the same three streams for the CPML psi."""
    N = cfg.npoints
    return w * (N * (3 * cfg.nv + n_background(cfg)) + 3 * cfg.npsi * psi_points(cfg))


def alg_bytes_zpass(cfg, w):
    """MHD z pass: read vz, p, bx, by, vx, vy and box, boy, boz; write the six
    z derivatives; the z-face psi streams."""
    return w * (cfg.npoints * (9 + 6) + 3 * cfg.npsi * psi_points(cfg, (4, 5)))


def alg_bytes_xy(cfg, w):
    """MHD xy kernel (z derivatives read from the z pass): the stage streams,
    the 3-D backgrounds, the six z-derivative arrays, the x/y-face psi streams."""
    N = cfg.npoints
    return w * (N * (3 * cfg.nv + n_background(cfg) + 6) + 3 * cfg.npsi * psi_points(cfg, (0, 1, 2, 3)))


def alg_bytes_step(cfg, w):
    """B/pt x N for a whole step: 5 stages + amortised filters (+ the Kahan
    term in the FP32 form: read and written once per step, 2 bytes per value
    -- the compensation is stored as bfloat16; a 4-byte term would add
    2 nv x 2 B/pt more)."""
    N = cfg.npoints
    nv = cfg.nv
    per_pt = 15 * nv + 5 * n_background(cfg) + 2 * nv * (1.0 / cfg.zfreq + 1.0 / cfg.xyfreq)
    b = w * (N * per_pt + 15 * cfg.npsi * psi_points(cfg))
    if cfg.precision == "fp32":
        b += 2 * nv * 2 * N
    return b


def profiled_traffic(workload, precision, tile, lane_points=1):
    """HBM bytes per launch of the stage kernel, from the committed rocprofv3
    summary profiles/<workload>_<precision>.md (FETCH_SIZE and WRITE_SIZE
    scaled by a 1 GiB copy), when it was taken with the same tile shape and
    points per lane."""
    path = os.path.join(ROOT, "profiles", f"{workload}_{precision}.md")
    if not os.path.exists(path):
        return None, None
    text = open(path).read()
    try:
        block = text.split("```json", 1)[1].split("```", 1)[0]
        summ = json.loads(block)
    except (IndexError, ValueError):
        return None, None
    if list(summ.get("tile") or []) != list(tile[:2]) or summ.get("lane_points", 1) != lane_points:
        return None, None
    # one RK stage: the stage kernel plus, for MHD, the z pass before it
    ks = [summ["kernels"].get(n) for n in ("k_stage", "k_zpass")]
    if not ks[0]:
        return None, None
    return sum(int(k["traffic_bytes_per_launch"]) for k in ks if k), os.path.relpath(path, ROOT)


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


def mode_precision(args):
    """FP32 fields + Kahan RK update against FP64 on quiet 512x512x300: the
    driver loop (steps + filters) on both, relative L2 of every field after
    500, 1000 and 2000 steps, and both throughputs."""
    from sparc2_amd import workloads
    from sparc2_amd.solver import Solver
    cfg64, bg = workloads.build("quiet_precision", precision="fp64")
    cfg32 = cfg64.replace(precision="fp32")
    s64 = Solver(cfg64, bg)
    s32 = Solver(cfg32, bg, setup=None)
    names = ["rho", "vx", "vy", "vz", "p"]
    marks = (500, 1000, 2000)
    l2, t = {}, {"fp64": 0.0, "fp32": 0.0}
    n0 = 0
    for m in marks:
        for key, s in (("fp64", s64), ("fp32", s32)):
            t0 = time.perf_counter()
            s.advance(n0, m - 1)
            s.sync()
            t[key] += time.perf_counter() - t0
        a64, a32 = s64.get_state(), s32.get_state()
        l2[m] = {f: float(np.linalg.norm(a32[q] - a64[q]) / max(np.linalg.norm(a64[q]), 1e-300))
                 for q, f in enumerate(names)}
        log(f"[precision] step {m}: {l2[m]}")
        n0 = m
    n = marks[-1]
    g64 = cfg64.npoints * n / t["fp64"] / 1e9
    g32 = cfg64.npoints * n / t["fp32"] / 1e9
    print(json.dumps({
        "metric": METRIC, "mode": "precision", "value": round(g32, 5), "unit": "GPt-updates/s",
        "n_gpus": 1, "steps": n, "higher_is_better": True, "dtype": "f32+kahan",
        "data": "synthetic (polytrope background, pulse source)",
        "config": {"workload": "quiet_precision", "grid": [cfg64.nx, cfg64.ny, cfg64.nz]},
        "fp64_gpts": round(g64, 5), "fp32_gpts": round(g32, 5), "speedup": round(g32 / g64, 3),
        "rel_l2": {str(k): v for k, v in l2.items()},
        "max_rel_l2": max(max(v.values()) for v in l2.values())}), flush=True)


def _batch_run(pop, precision, warmup, steps, autotune):
    """Advance the population batched in one precision; returns GPt/s,
    ms/step, the tuned groups, launch sizes and the runs that failed the
    reference's checks (their points are not counted)."""
    from sparc2_amd.batch import Batch
    from sparc2_amd.solver import Solver
    t0 = time.perf_counter()
    sol = [Solver(c.replace(precision=precision), b) for c, b in pop]
    log(f"[batch {precision}] {len(sol)} runs set up in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    b = Batch(sol, use_graph=True, autotune=autotune)
    log(f"[batch {precision}] batch created (autotune {autotune}) in {time.perf_counter() - t0:.1f} s")
    b.advance(warmup, check_guard=False)
    for s in sol:
        s.sync()
    t0 = time.perf_counter()
    b.advance(steps, check_guard=False)
    for s in sol:
        s.sync()
    el = time.perf_counter() - t0
    failed = b.failures()
    ok = [q for q in range(len(pop)) if q not in failed]
    pts = sum(pop[q][0].npoints for q in ok)
    w = 4 if precision == "fp32" else 8
    alg = sum(alg_bytes_step(pop[q][0].replace(precision=precision), w) for q in ok)
    groups, blocks = b.info()
    tiles = ([[int(k[0]), int(k[1]), *v] for k, v in sorted(b.tiles.items())] if b.tiles else None)
    b.close()
    for s in sol:
        s.close()
    if failed:
        log(f"[batch {precision}] runs stopped by the failure checks (not counted): "
            f"{ {q: str(e) for q, e in failed.items()} }")
    return {"gpts": pts * steps / el / 1e9, "ms_per_step": el / steps * 1e3,
            "alg_GBps": alg * steps / el / 1e9, "groups": groups, "launch_blocks": blocks,
            "tuned_groups": tiles, "failed_runs": sorted(failed)}


def mode_batch(args):
    """The batched form: the seeded population of independent runs advanced
    together in FP32 with the compensated update, and the same population
    batched in plain FP64 (its twin); aggregate point-updates per second and
    the algorithmic HBM fraction of each."""
    from sparc2_amd.workloads import batch_population
    pop = batch_population(args.runs or 64)
    K = args.steps
    f32 = _batch_run(pop, "fp32", args.warmup, K, not args.no_autotune)
    f64 = _batch_run(pop, "fp64", args.warmup, K, not args.no_autotune)
    ceiling = stream_ceiling()
    for r in (f32, f64):
        r["hbm_frac"] = round(r["alg_GBps"] / HBM_PEAK_GBS, 4)
        r["frac_of_stream_ceiling_4:1"] = round(r["alg_GBps"] / ceiling["4:1"], 4)
    n_mhd = sum(1 for c, _ in pop if c.mhd)
    print(json.dumps({
        "metric": METRIC, "mode": "batch", "value": round(f32["gpts"], 5),
        "unit": "GPt-updates/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": round(f32["ms_per_step"], 4), "higher_is_better": True,
        "dtype": "f32+kahan",
        "data": "synthetic (seeded population: polytrope / flux-tube backgrounds)",
        "config": {"workload": "batch_population", "runs": len(pop), "mhd_runs": n_mhd,
                   "points": sum(c.npoints for c, _ in pop), "graph": True,
                   "autotune": not args.no_autotune},
        "roofline": {"bound": "hbm", "achieved": round(f32["alg_GBps"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": f32["hbm_frac"],
                     "frac_of_stream_ceiling_4:1": f32["frac_of_stream_ceiling_4:1"],
                     "stream_ceiling_GBps": ceiling,
                     "basis": "algorithmic bytes per step of every counted run x steps / time"},
        "fp32": {k: (round(v, 5) if isinstance(v, float) else v) for k, v in f32.items()},
        "fp64_twin": {k: (round(v, 5) if isinstance(v, float) else v) for k, v in f64.items()},
        "speedup_vs_fp64_batched": round(f32["gpts"] / f64["gpts"], 3)}), flush=True)


def mode_serve(args):
    """The served form: a seeded Poisson stream of requests through the
    continuous-batching server, in each precision of --serve-precisions;
    runs/h, throughput and latency percentiles of each.  --step-scale
    shrinks the drawn step counts (a bounded sample of the stream)."""
    from sparc2_amd.serve import Server, poisson_stream
    n = args.runs or 8
    lo, hi = (max(1, int(round(v * args.step_scale))) for v in (500, 5000))
    out = {}
    for prec in args.serve_precisions.split(","):
        reqs = poisson_stream(n, precision=prec, steps=(lo, hi))
        r = Server(autotune=not args.no_autotune).run(reqs, verbose=True)
        failed = [rid for rid, res in r["results"].items() if res.error is not None]
        out[prec] = {"runs_per_h": round(r["runs_per_h"], 2), "gpts": round(r["gpts"], 5),
                     "gpts_busy": round(r["gpts_busy"], 5), "wall_s": round(r["wall_s"], 3),
                     "rebuild_s": round(r["rebuild_s"], 3),
                     "latency_p50_s": round(r["latency_p50"], 3),
                     "latency_p95_s": round(r["latency_p95"], 3),
                     "latency_p99_s": round(r["latency_p99"], 3), "failed_requests": failed}
        log(f"[serve] {prec}: {out[prec]}")
    first = list(out)[0]
    line = {
        "metric": METRIC, "mode": "serve", "value": out[first]["runs_per_h"], "unit": "runs/h",
        "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32+kahan" if first == "fp32" else "f64",
        "data": "synthetic (seeded Poisson stream, rate 4/s)",
        "config": {"workload": "poisson_stream", "requests": n, "steps_drawn_from": [lo, hi],
                   "grids": "64^3 .. 384x384x256", "autotune": not args.no_autotune}}
    line.update(out)
    if "fp32" in out and "fp64" in out:
        line["gpts_ratio_fp32_fp64"] = round(out["fp32"]["gpts_busy"] / out["fp64"]["gpts_busy"], 3)
    print(json.dumps(line), flush=True)


def mode_profile(args):
    """What rocprofv3 runs: one device copy of 1 GiB and two launches of the
    library's stream kernel (4 arrays of 512 MiB read, one written) -- their
    FETCH_SIZE / WRITE_SIZE readings calibrate the counters against known
    byte counts of two access patterns -- then --steps driver steps of the
    workload.  No timing of its own."""
    import ctypes as C

    import torch
    from sparc2_amd._lib import check, lib
    a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    torch.cuda.empty_cache()
    g = C.c_double()
    check(lib().sparc_hip_stream_bench(0, 1 << 26, 4, 2, C.byref(g)), "sparc_hip_stream_bench")
    from sparc2_amd import workloads
    from sparc2_amd.solver import Solver
    cfg, bg = workloads.build(args.workload, precision=args.precision)
    tile = tuple(int(v) for v in args.tile.split(","))
    s = Solver(cfg, bg, tile=tile)
    s.advance(1, args.steps)
    s.sync()
    print(json.dumps({"mode": "profile", "workload": args.workload, "precision": args.precision,
                      "steps": args.steps, "tile": list(s.tile()[0]), "copy_bytes": 1 << 30,
                      "stream_read_bytes": 4 << 29, "stream_write_bytes": 1 << 29}),
          flush=True)
    s.close()


def stream_ceiling(n=1 << 26, reps=20, reads=(1, 2, 4, 8)):
    """Measured bandwidth of the library's own stream probe (nread arrays
    summed into one, 16-byte non-temporal accesses) at several read-to-write
    ratios: the ceiling the stage kernels' fractions are also read against."""
    import ctypes as C
    from sparc2_amd._lib import check, lib
    out = {}
    for r in reads:
        g = C.c_double()
        check(lib().sparc_hip_stream_bench(0, n, r, reps, C.byref(g)), "sparc_hip_stream_bench")
        out[f"{r}:1"] = round(g.value, 1)
    log(f"[stream] GB/s by reads:writes {out}")
    return out


def precision_form(marks=(500, 1000, 2000)):
    """The precision form on quiet 512x512x300: FP32 fields with the
    compensated RK update against plain FP64, the same driver loop on both;
    relative L2 of every field at each mark, throughput of each over the
    whole run (state resident; set-up and read-outs excluded)."""
    from sparc2_amd import workloads
    from sparc2_amd.solver import Solver
    cfg64, bg = workloads.build("quiet_precision", precision="fp64")
    t0 = time.perf_counter()
    s64 = Solver(cfg64, bg)
    cfg32 = cfg64.replace(precision="fp32")
    s32 = Solver(cfg32, bg, setup=s64.st)
    log(f"[precision] set-up {time.perf_counter() - t0:.1f} s")
    names = ["rho", "vx", "vy", "vz", "p"]
    l2, t = {}, {"fp64": 0.0, "fp32": 0.0}
    n0 = 0
    for m in marks:
        for key, s in (("fp64", s64), ("fp32", s32)):
            s.sync()
            t1 = time.perf_counter()
            s.advance(n0, m - 1)        # ends with the guard read (one sync)
            t[key] += time.perf_counter() - t1
        a64, a32 = s64.get_state(), s32.get_state()
        l2[str(m)] = {f: float(np.linalg.norm(a32[q] - a64[q]) / max(np.linalg.norm(a64[q]), 1e-300))
                      for q, f in enumerate(names)}
        log(f"[precision] step {m}: {l2[str(m)]}")
        n0 = m
    n = marks[-1]
    g64 = cfg64.npoints * n / t["fp64"] / 1e9
    g32 = cfg64.npoints * n / t["fp32"] / 1e9
    b64, b32 = alg_bytes_step(cfg64, 8) / cfg64.npoints, alg_bytes_step(cfg32, 4) / cfg64.npoints
    s64.close()
    s32.close()
    return {"workload": "quiet_precision", "grid": [cfg64.nx, cfg64.ny, cfg64.nz], "steps": n,
            "fp64_gpts": round(g64, 4), "fp32_gpts": round(g32, 4), "speedup": round(g32 / g64, 3),
            "alg_bytes_per_point": {"fp64": round(b64, 1), "fp32": round(b32, 1)},
            "hbm_frac": {"fp64": round(b64 * g64 / HBM_PEAK_GBS, 4),
                         "fp32": round(b32 * g32 / HBM_PEAK_GBS, 4)},
            "rel_l2": l2, "max_rel_l2": max(max(v.values()) for v in l2.values())}


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
    ap.add_argument("--no-precision", action="store_true",
                    help="leave out the FP32-vs-FP64 comparison on quiet 512x512x300")
    ap.add_argument("--tile", default="0,0,0")
    ap.add_argument("--mode", default="run", choices=["run", "precision", "batch", "serve", "profile"],
                    help="run: the headline single-run line (default); precision: FP32+Kahan "
                         "vs FP64 relative L2 on quiet 512x512x300; batch: 64 seeded runs in "
                         "one batch; serve: a Poisson request stream; profile: a 1 GiB copy "
                         "(the HBM-counter calibration) then --steps driver steps, for rocprofv3")
    ap.add_argument("--no-autotune", action="store_true", help="batch: keep the default tiles")
    ap.add_argument("--serve-precisions", default="fp32,fp64",
                    help="serve: precisions to run the stream in, comma separated")
    ap.add_argument("--step-scale", type=float, default=1.0,
                    help="serve: scale of the drawn step counts (500..5000 at 1)")
    ap.add_argument("--runs", type=int, default=None,
                    help="batch: number of runs (64); serve: number of requests (8)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: the whole domain lives on one device; "
              "multi-GPU sharding is not built", file=sys.stderr)
        sys.exit(2)
    if args.mode != "run":
        {"precision": mode_precision, "batch": mode_batch, "serve": mode_serve,
         "profile": mode_profile}[args.mode](args)
        return

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
    s.advance(W, W + K - 1, check_guard=False)
    s.sync()
    elapsed = time.perf_counter() - t0
    ms_step = elapsed / K * 1e3
    gpts = cfg.npoints * K / elapsed / 1e9
    # the timed steps passed the reference's failure checks (a run stopped by
    # them would skip its kernels and time nothing)
    s.check_guard()
    d = s.diag()
    log(f"[bench] {K} steps in {elapsed:.3f} s: {ms_step:.3f} ms/step, {gpts:.3f} GPt/s, "
        f"diag {d}")

    # stage kernels timed with HIP events on the context's stream
    P = args.profile_steps
    s.timing(True)
    s.advance(W + K, W + K + P - 1)
    s.sync()
    times = {}
    for kind, name in ((0, "stage1"), (1, "stage2_4"), (2, "stage5"), (3, "filter"), (4, "zpass")):
        n, ms = s.kernel_time(kind)
        times[name] = (n, ms / n if n else None)
    s.timing(False)
    n_mid, avg_mid = times["stage2_4"]
    bytes_stage = alg_bytes_stage(cfg, w)
    zpass = times["zpass"][0] > 0
    bytes_xy = alg_bytes_xy(cfg, w) if zpass else bytes_stage
    per_kernel = {"k_stage": {"avg_ms": round(avg_mid, 5), "alg_bytes": int(bytes_xy),
                              "frac": round(bytes_xy / (avg_mid * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}
    stage_ms = avg_mid
    split_extra = 0
    if zpass:
        zb = alg_bytes_zpass(cfg, w)
        zms = times["zpass"][1]
        per_kernel["k_zpass"] = {"avg_ms": round(zms, 5), "alg_bytes": int(zb),
                                 "frac": round(zb / (zms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        stage_ms = avg_mid + zms
        # the z derivatives handed from the z pass to the xy kernel (written,
        # read back) and the z pass's own reads: bytes the split adds
        split_extra = int(bytes_xy + zb - bytes_stage)
    # graded figure: one middle RK stage against its unsplit algorithmic bytes
    achieved = bytes_stage / (stage_ms * 1e-3) / 1e9
    log(f"[bench] RK stage {stage_ms:.4f} ms, {bytes_stage / 1e9:.3f} GB algorithmic -> "
        f"{achieved:.1f} GB/s = {achieved / HBM_PEAK_GBS:.4f} of peak; kernels {times}")
    step_bytes = alg_bytes_step(cfg, w)
    tile_used, nblk = s.tile()
    # the library gives FP32 lanes two x points on grids with even nx
    lanes = 2 if (cfg.precision == "fp32" and cfg.nx % 2 == 0
                  and os.environ.get("SPARC_LANE_POINTS", "") != "1") else 1
    traffic, traffic_src = profiled_traffic(args.workload, args.precision, tile_used, lanes)
    s.close()

    ceiling = stream_ceiling()
    # the stage's read-to-write mix: MHD (u, a, 16 backgrounds) : un per
    # field ~ 4:1, quiet ~ 2:1
    ceil_key = "4:1" if cfg.mhd else "2:1"
    prec = None
    if not args.no_precision:
        prec = precision_form()
        for key in ("fp64", "fp32"):
            prec["hbm_frac"][f"{key}_vs_stream_2:1"] = round(
                prec["alg_bytes_per_point"][key] * prec[f"{key}_gpts"] / ceiling["2:1"], 4)

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
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": ("one middle RK stage = k_zpass + k_stage (MHD xy kernel); achieved = the "
                       "stage's unsplit algorithmic bytes / (z pass + xy kernel time)"
                       if zpass else "k_stage, RK stages 2-4"),
            "alg_bytes": int(bytes_stage),
            "stage_ms": round(stage_ms, 5),
            "split_extra_alg_bytes": split_extra,
            "frac_of_stream_ceiling": round(achieved / ceiling[ceil_key], 4),
            "stream_ceiling_GBps": ceiling,
            "per_kernel": per_kernel,
        },
        "precision_form": prec,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
