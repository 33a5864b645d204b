"""Served form: forward-simulation requests with continuous batching.

A request is one SPARC2 run (grid, background, source position, step count).
The server admits arrived requests in arrival order while their device
footprint fits the HBM budget, keeps every admitted run's state resident, and
advances all admitted runs together with the batched kernels.  Between
batched advances it returns each run's photospheric vz slice every 45 s of
simulated time (the full-width chunk output of WRITE_OUT_CHUNK,
all_modules.f90:1566-1799, at the plane nearest z = 0), retires finished runs
and admits waiting ones -- requests that arrive while others are running join
at the next advance.  Latency is measured from arrival to completion.
"""
from __future__ import annotations

import dataclasses
import math
import os
import sys
import time
from typing import Callable, Dict, List, Optional

import numpy as np

from .background import make_background
from .config import RunConfig
from .initialise import prepare

VZ = 3


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float                 # s after the stream start
    cfg: RunConfig
    zrange: tuple                  # (zbot, ztop) in Mm
    nsteps: int
    B0: float = 0.0                # G; 0 = quiet
    slice_every_s: float = 45.0

    def build_background(self):
        if self.cfg.magnetic == 2:
            return make_background(self.cfg, *self.zrange, B0=self.B0, R0=4.0e8)
        return make_background(self.cfg, *self.zrange)


@dataclasses.dataclass
class Result:
    rid: int
    arrival: float
    admitted: float = math.nan
    finished: float = math.nan
    slices: List = dataclasses.field(default_factory=list)   # (time index, 2-D vz)
    final: Optional[np.ndarray] = None
    points: int = 0
    steps: int = 0
    # set when the run failed the reference's checks (driver.f90:175-196);
    # it is retired at the failing step
    error: Optional[Exception] = None

    @property
    def latency(self):
        return self.finished - self.arrival


def device_footprint(cfg: RunConfig) -> int:
    """Bytes a run keeps on the device: 3 state buffers (+2 bfloat16
    compensation buffers in FP32, each half a state buffer), backgrounds,
    CPML arrays."""
    w = 4 if cfg.precision == "fp32" else 8
    fs = -(-cfg.npoints // 256) * 256
    n_state = 3 + (1 if cfg.precision == "fp32" else 0)
    n_bg = 18 if cfg.mhd else 7
    psi = sum(int(np.prod(cfg.face_shape(f))) for f in range(6) if cfg.face_active(f))
    return w * (n_state * cfg.nv * fs + n_bg * fs + psi * (3 + 2 * cfg.npsi)) + (1 << 20)


def device_memory(device: int = 0):
    """(free, total) bytes of the device (hipMemGetInfo)."""
    import ctypes as C
    from ._lib import check, lib
    f, t = C.c_long(), C.c_long()
    check(lib().sparc_hip_device_memory(int(device), C.byref(f), C.byref(t)), "sparc_hip_device_memory")
    return f.value, t.value


def surface_plane(zaxis_mm) -> int:
    return int(np.argmin(np.abs(np.asarray(zaxis_mm))))


def cfl_timestep(cfg: RunConfig, bg, vmax_alfven=1.0e7, courant=1.0, quantum=0.05) -> float:
    """Time step from the reference's CFL estimate (driver.f90:83-117): the
    smallest vertical crossing time of the fastest wave, with the Alfven speed
    capped by the limiter, rounded down to `quantum` seconds."""
    z = np.asarray(bg["zaxis"]) * 1e8
    dz = np.diff(z)
    c = np.asarray(bg["cs0"])
    vmax = c.max(axis=(1, 2))
    if cfg.magnetic == 2:
        vmax = np.sqrt(vmax ** 2 + vmax_alfven ** 2)
    t = courant * float(np.min(dz / np.maximum(vmax[:-1], vmax[1:])))
    return max(quantum, math.floor(t / quantum) * quantum)


def poisson_stream(n: int, rate: float = 4.0, seed: int = 7, precision: str = "fp32",
                   grids=((64, 64, 64), (128, 128, 128), (192, 192, 160), (256, 256, 192),
                          (384, 384, 256)),
                   steps=(500, 5000), pulse_amp: float = 1.0e2) -> List[Request]:
    """Seeded request stream: Poisson arrivals, grids drawn from `grids`,
    B0 uniform in [0, 3] kG (below 1 G the quiet kernel), random source
    position in the central half, step count uniform in `steps`.  The pulse
    is 1 m/s by default (the equations are linear; at 100 m/s the longer
    flux-tube runs drive p + p0 negative in their tenuous top layers, which
    the reference's failure checks stop)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n):
        t += rng.exponential(1.0 / rate)
        nx, ny, nz = grids[rng.integers(len(grids))]
        B0 = float(rng.uniform(0.0, 3000.0))
        mhd = B0 >= 1.0
        L = 100.0e8 if mhd else 50.0e8
        px, py = rng.uniform(0.25 * L, 0.75 * L, size=2)
        zr = (-20.0, 1.0) if mhd else (-10.0, 0.5)
        nsteps = int(rng.integers(steps[0], steps[1] + 1))
        cfg = RunConfig(
            nx=nx, ny=ny, nz=nz, xlength=L, ylength=L, timestep=1.0, magnetic=2 if mhd else 0,
            periodic=not mhd, use_pml=True, use_hpml=mhd,
            npml=(10, 10, 10, 10, 10, 10) if mhd else (0, 0, 0, 0, 10, 10),
            pmlc=((9.2103, 2.0, 8.0, 0.1),) * 6, t_xy=30.0, t_z=10.0,
            pulse_t0=1000.0 / 3.0, pulse_t1=450.0, pulse_st=150.0,
            pulse_px=float(px), pulse_py=float(py), pulse_pz=-0.15e8,
            pulse_sr=1.0e8, pulse_sz=1.0e8, pulse_amp=pulse_amp, maxva=1.0e7, precision=precision)
        out.append(Request(rid=rid, arrival=t, cfg=cfg, zrange=zr, nsteps=nsteps, B0=B0))
    return out


class Server:
    """Continuous-batching server.  A request's set-up -- background, CFL
    time step, FP64 set-up and its device context -- runs on a worker thread
    (``setup_workers`` of them) while the admitted runs advance; its HBM
    footprint is reserved when its set-up starts, in arrival order, and it
    joins the batch at the first advance after its set-up has finished.  While
    set-ups are in flight the advances are cut to ``poll_chunk`` steps so a
    ready run joins promptly.  With ``chunk_dir`` each returned slice is also
    written there as a chunk file of the surface plane (all fields, the
    layout of WRITE_OUT_CHUNK, all_modules.f90:1566-1799), named
    ``req<rid>_chunk_NNNN.h5`` with NNNN the slice count."""

    def __init__(self, hbm_budget: Optional[float] = None, max_chunk: int = 200, use_graph: bool = True,
                 device: int = 0, keep_final: bool = False, setup_workers: int = 2,
                 poll_chunk: int = 25, chunk_dir: Optional[str] = None,
                 autotune: bool = False,
                 clock: Callable[[], float] = time.perf_counter):
        # None: 90% of the device memory free when the stream starts
        self.budget = hbm_budget
        self.max_chunk = max_chunk
        self.use_graph = use_graph
        self.device = device
        self.keep_final = keep_final
        self.setup_workers = max(1, setup_workers)
        self.poll_chunk = max(1, poll_chunk)
        self.chunk_dir = chunk_dir
        # autotune: tile and z chunk tuned once per kernel variant (the first
        # batch that holds it) and reused by every later membership it fits
        # (batch.tune_tiles with by_variant)
        self.autotune = autotune
        self.clock = clock

    def _start(self, req: Request):
        from .solver import Solver
        bg = req.build_background()
        cfg = req.cfg
        cfg.timestep = cfl_timestep(cfg, bg)
        s = Solver(cfg, bg, device=self.device)
        every = max(1, int(round(req.slice_every_s / cfg.timestep)))
        return {"solver": s, "done": 0, "every": every, "k0": surface_plane(bg["zaxis"])}

    def run(self, requests: List[Request], verbose: bool = False) -> Dict:
        from concurrent.futures import ThreadPoolExecutor
        from .batch import Batch
        if self.budget is None:
            free, _ = device_memory(self.device)
            self.budget = 0.9 * free
        reqs = sorted(requests, key=lambda r: r.arrival)
        results = {r.rid: Result(rid=r.rid, arrival=r.arrival) for r in reqs}
        waiting: List[Request] = []    # arrived, set-up not started
        pending: List[tuple] = []      # (request, footprint, future): set-up running
        active: Dict[int, dict] = {}
        used = 0.0
        t0 = self.clock()
        now = lambda: self.clock() - t0
        nxt = 0
        batch, members = None, ()
        busy = 0.0
        rebuild = 0.0     # batch (re)creation and tuning when membership changes
        say = (lambda m: print(m, file=sys.stderr, flush=True)) if verbose else (lambda m: None)
        pool = ThreadPoolExecutor(max_workers=self.setup_workers)
        try:
            while nxt < len(reqs) or waiting or pending or active:
                while nxt < len(reqs) and reqs[nxt].arrival <= now():
                    waiting.append(reqs[nxt])
                    nxt += 1
                # reserve HBM and start set-ups in arrival order while the footprint fits
                while waiting:
                    fp = device_footprint(waiting[0].cfg)
                    if (active or pending) and used + fp > self.budget:
                        break
                    req = waiting.pop(0)
                    used += fp
                    pending.append((req, fp, pool.submit(self._start, req)))
                # nothing on the device: wait for the oldest set-up
                if not active and pending:
                    pending[0][2].result()
                still = []
                for req, fp, fut in pending:
                    if not fut.done():
                        still.append((req, fp, fut))
                        continue
                    st = fut.result()
                    st["req"], st["fp"] = req, fp
                    active[req.rid] = st
                    results[req.rid].admitted = now()
                    results[req.rid].points = req.cfg.npoints
                    say(f"[serve] t={now():7.2f}s admit {req.rid} {req.cfg.nx}x{req.cfg.ny}x{req.cfg.nz} "
                        f"{'MHD' if req.cfg.mhd else 'quiet'} {req.nsteps} steps")
                pending = still
                if not active:
                    if nxt < len(reqs):
                        time.sleep(max(0.0, min(0.05, reqs[nxt].arrival - now())))
                    continue
                key = tuple(sorted(active))
                if key != members:
                    if batch is not None:
                        batch.filter_deferred()
                        batch.close()
                    sol = [active[r]["solver"] for r in key]
                    tb = self.clock()
                    batch = Batch(sol, use_graph=self.use_graph, autotune=self.autotune, tune_by_variant=True)
                    rebuild += self.clock() - tb
                    batch.time = [active[r]["done"] for r in key]
                    members = key
                # advance to the next event of any run: a slice or its end
                chunk = self.poll_chunk if (pending or waiting) else self.max_chunk
                for r in key:
                    a = active[r]
                    left = a["req"].nsteps - a["done"]
                    to_slice = a["every"] - (a["done"] % a["every"])
                    chunk = min(chunk, left, to_slice)
                tb = self.clock()
                # the slices are the reference's chunk output: the state after
                # STEP and before that step's FILTER_VARS (driver.f90:198-200)
                failed = batch.advance(chunk, defer_filter=True)
                busy += self.clock() - tb
                for q, r in enumerate(key):
                    a = active[r]
                    a["done"] += chunk
                    res = results[r]
                    res.steps = a["done"]
                    if q in failed:
                        # stopped at the failing step; retired below, once the
                        # batch has no more work queued for it
                        res.error = failed[q]
                        res.steps = failed[q].time + 1
                        say(f"[serve] t={now():7.2f}s fail  {r}: {failed[q]}")
                        continue
                    if a["done"] % a["every"] == 0:
                        res.slices.append((a["done"], a["solver"].get_slab(VZ, a["k0"], a["k0"])[0]))
                        if self.chunk_dir:
                            from . import h5io
                            name = h5io.file_name(f"req{r}_", "chunk", len(res.slices))
                            h5io.write_chunk(os.path.join(self.chunk_dir, name), a["solver"], a["done"],
                                             zr=(a["k0"], a["k0"]))
                batch.filter_deferred()
                for r in key:
                    a = active.get(r)
                    if a is None:
                        continue
                    res = results[r]
                    if res.error is not None or a["done"] >= a["req"].nsteps:
                        if self.keep_final and res.error is None:
                            res.final = a["solver"].get_state()
                        res.finished = now()
                        a["solver"].close()
                        used -= a["fp"]
                        del active[r]
                        say(f"[serve] t={now():7.2f}s done  {r} latency {res.latency:.2f}s")
        finally:
            pool.shutdown(wait=True)
            if batch is not None:
                batch.close()
        wall = now()
        lat = np.array([results[r.rid].latency for r in reqs])
        pts = sum(results[r.rid].points * results[r.rid].steps for r in reqs)
        return {
            "results": results,
            "wall_s": wall,
            "busy_s": busy,
            "rebuild_s": rebuild,
            "runs_per_h": len(reqs) / wall * 3600.0,
            "gpts": pts / wall / 1e9,
            "gpts_busy": pts / busy / 1e9 if busy > 0 else float("nan"),
            "latency_p50": float(np.percentile(lat, 50)),
            "latency_p95": float(np.percentile(lat, 95)),
            "latency_p99": float(np.percentile(lat, 99)),
        }
