"""The reference's time loop on one resident run (driver.f90:123-202).

    run_driver(solver, n0, n1, out_dir="out", simname="run_",
               fulloutputsteps=600, chunkoutputsteps=90)

Each time index nn = n0..n1 does, in the reference's order: the source of
step nn, STEP, the failure checks (norm_KE > 1e3, min(rho+rho0) <= 0,
min(p+p0) <= 0; on failure the state is written to full_XXXX.h5 and the
run stops), the full snapshot full_NNNN.h5 every ``fulloutputsteps`` steps
and at time 0, the chunk file chunk_NNNN.h5 every ``chunkoutputsteps`` steps
once ``minsavtime`` hours have passed and after the first step of the loop,
and FILTER_VARS when nn != 0.  Outputs therefore hold the state after STEP
and before that step's filter, as the reference writes them.

Between output steps the run advances asynchronously; the device evaluates
the failure checks of every step and halts the run at the first failure, so
one synchronisation per output (or per ``sync_every`` steps) suffices.

Restart (driver.f90:61-69): ``restart(solver, path, fulloutputsteps)`` loads
full_NNNN.h5 and returns the next time index, time0 * fulloutputsteps + 1;
the snapshot's own step is not filtered again, as in the reference.
"""
from __future__ import annotations

import math
import os
from typing import Callable, Dict, List, Optional

from . import h5io


def _full_name(simname: str, time: int, fulloutputsteps: int) -> str:
    """full_NNNN.h5 with NNNN = time / fulloutputsteps (0 at time 0),
    all_modules.f90:1088-1098."""
    n = 0 if time == 0 else time // max(1, fulloutputsteps)
    return h5io.file_name(simname, "full", n)


def run_driver(solver, n0: int, n1: int, out_dir: Optional[str] = None, simname: str = "",
               fulloutputsteps: int = 0, chunkoutputsteps: int = 0, minsavtime: float = 0.0,
               chunk_region: Optional[Dict] = None, sync_every: int = 0,
               on_output: Optional[Callable[[str, int], None]] = None) -> List[tuple]:
    """The time loop nn = n0..n1 of driver.f90:123-202.  fulloutputsteps or
    chunkoutputsteps 0 switches that output off (FULLOUT / CHUNKOUT false);
    the full snapshot at time 0 is written regardless, as the reference does.
    chunk_region: optional {"xr": (i0, i1), "yr": ..., "zr": ...} (0-based,
    inclusive) of the chunk files.  on_output(kind, time) is called after
    each output ("full" or "chunk") -- also without out_dir, for callers
    that read the state themselves.  Returns [(kind, time, path)].
    Raises DivergenceError (after writing full_XXXX.h5 to out_dir) when a
    step fails the reference's checks."""
    cfg = solver.cfg
    written: List[tuple] = []
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
    init = n0

    def full_due(t):
        return (fulloutputsteps > 0 and t % fulloutputsteps == 0) or t == 0

    def chunk_due(t):
        return (chunkoutputsteps > 0 and t % chunkoutputsteps == 0
                and t * cfg.timestep >= minsavtime * 3600.0 and t > init)

    def next_event(t):
        """First time index >= t with an output (or a forced sync), else n1."""
        cands = [n1]
        if t <= 0 <= n1:
            cands.append(0)
        for every in (fulloutputsteps, chunkoutputsteps, sync_every):
            if every > 0:
                cands.append(int(math.ceil(t / every)) * every)
        cands = [c for c in cands if t <= c <= n1]
        for c in sorted(cands):
            if full_due(c) or chunk_due(c) or c == n1 or (sync_every and c % sync_every == 0):
                return c
        return n1

    t = n0
    while t <= n1:
        e = next_event(t)
        # steps t..e-1 with their filters, step e without its filter
        solver.advance(t, e, check_guard=False, filter_last=False)
        solver.check_guard(out_dir, simname)
        if full_due(e):
            path = None
            if out_dir:
                path = os.path.join(out_dir, _full_name(simname, e, fulloutputsteps))
                h5io.write_full(path, solver, time=e)
            written.append(("full", e, path))
            if on_output:
                on_output("full", e)
        if chunk_due(e):
            path = None
            if out_dir:
                n = e // chunkoutputsteps
                path = os.path.join(out_dir, h5io.file_name(simname, "chunk", n))
                h5io.write_chunk(path, solver, e, **(chunk_region or {}))
            written.append(("chunk", e, path))
            if on_output:
                on_output("chunk", e)
        if e != 0:
            solver.filter_vars(e)
        t = e + 1
    return written


def restart(solver, path: str, fulloutputsteps: int) -> int:
    """READ_IN_FULL_STATE of full_NNNN.h5 (all_modules.f90:741-1069): loads
    the state and CPML bases and returns the time index to continue from,
    time0 * fulloutputsteps + 1 (driver.f90:64-65), where time0 is the
    snapshot number NNNN taken from the file name."""
    h5io.read_full(path, solver)
    stem = os.path.basename(path)
    num = stem.rsplit("_", 1)[-1].split(".", 1)[0]
    if not num.isdigit():
        raise ValueError(f"{path}: no snapshot number in the name")
    return int(num) * fulloutputsteps + 1
