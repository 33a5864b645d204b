"""Batched form: many independent runs advanced by shared persistent launches.

    b = Batch([Solver(cfg_i, bg_i) for ...], use_graph=True)
    b.advance(1000)          # every run advances 1000 driver steps

Each run keeps its own time index; its source amplitudes (driver.f90:140-142)
and filter cadence (derivatives.f90:669,690) follow that index.  The runs'
states stay device-resident in their own contexts; ``Solver.get_state`` and
``Solver.diag`` read them as usual.  Each run's failure checks
(driver.f90:175-196) run on the device every step; a run that fails stops
at that step (its later work items are skipped) and ``advance`` reports it
without disturbing the others.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, List, Sequence

import numpy as np

from ._lib import DEFER_FILTER, SparcError, check, dptr, lib
from .solver import DivergenceError, Solver


# Tile shapes the stage kernels are built for.
TILES = ((16, 16), (32, 8), (32, 16), (64, 4), (64, 8))


def _create(solvers: Sequence[Solver], use_graph: bool):
    n = len(solvers)
    hs = (C.c_void_p * n)(*[s.h.value for s in solvers])
    h = C.c_void_p()
    check(lib().sparc_hip_batch_create(hs, n, int(use_graph), C.byref(h)), "sparc_hip_batch_create")
    return h


def _time_stage(h, reps: int) -> float:
    ms = C.c_double()
    check(lib().sparc_hip_batch_time_stage(h, int(reps), C.byref(ms)), "sparc_hip_batch_time_stage")
    return ms.value


# z-chunk depths the batch tuner tries, as multiples of each run's automatic
# chunk: a persistent block pays a fixed cost per work item (its staging
# plan, the argument switch, a barrier), which deeper chunks spread over more
# planes.
KC_MULT = (1, 2, 4)


def _apply(members: Sequence[Solver], tile, auto, mult):
    for s, k in zip(members, auto):
        s.set_tile((tile[0], tile[1], 0 if mult == 1 else min(k * mult, s.cfg.nz)))


# Tuned choices: keyed by the group's variant and the sorted tile keys of
# its runs (a group that comes back is not timed again), or with by_variant
# by the variant, precision and lane setting alone -- the served form, whose
# membership changes at every arrival and departure, times a variant once
# and reuses its shape for every later group of that variant that it fits.
_GROUP_CACHE: Dict[tuple, tuple] = {}


def tune_tiles(solvers: Sequence[Solver], reps: int = 3, by_variant: bool = False) -> Dict[tuple, tuple]:
    """Tile autotuning for a batch.  Runs that share a kernel variant share
    one launch, so the shape is chosen per variant group, by timing a middle
    RK stage of that group's batched launch for every shape that fits all of
    its grids, each at the runs' automatic z chunks and at KC_MULT times
    them (timing leaves the runs' states unchanged).  Returns
    {variant: (tx, ty, chunk multiple)} and leaves each run on its group's
    best choice."""
    groups: Dict[tuple, List[Solver]] = {}
    for s in solvers:
        groups.setdefault((s.cfg.magnetic == 2, bool(s.st.bg3d)), []).append(s)
    chosen = {}
    for key, members in groups.items():
        if by_variant:
            tk = members[0].tile_key()
            ckey = ("variant", key, tk[4], tk[-1])
        else:
            ckey = (key, tuple(sorted(s.tile_key() for s in members)))
        if ckey in _GROUP_CACHE:
            t, mult = _GROUP_CACHE[ckey]
            try:
                for s in members:
                    s.set_tile((t[0], t[1], 0))
            except SparcError:
                pass                    # the shape does not fit a new grid: time again
            else:
                auto = [s.tile()[0][2] for s in members]
                _apply(members, t, auto, mult)
                chosen[key] = (t[0], t[1], mult)
                continue
        best, best_ms, best_auto = None, float("inf"), None
        for t in TILES:
            try:
                for s in members:
                    s.set_tile((t[0], t[1], 0))
            except SparcError:
                continue
            auto = [s.tile()[0][2] for s in members]
            for mult in KC_MULT:
                if mult > 1 and all(k * mult > s.cfg.nz for s, k in zip(members, auto)):
                    break
                _apply(members, t, auto, mult)
                h = _create(members, use_graph=False)
                try:
                    ms = _time_stage(h, reps)
                finally:
                    lib().sparc_hip_batch_destroy(h)
                if ms < best_ms:
                    best, best_ms, best_auto = (t[0], t[1], mult), ms, auto
        if best is None:
            raise SparcError(f"no tile shape fits every grid of variant {key}")
        _apply(members, best[:2], best_auto, best[2])
        chosen[key] = best
        _GROUP_CACHE[ckey] = (best[:2], best[2])
    return chosen


class Batch:
    def __init__(self, solvers: Sequence[Solver], use_graph: bool = True, autotune: bool = False,
                 tune_by_variant: bool = False):
        self.solvers: List[Solver] = list(solvers)
        if not self.solvers:
            raise ValueError("empty batch")
        self.tiles = tune_tiles(self.solvers, by_variant=tune_by_variant) if autotune else None
        self.h = _create(self.solvers, use_graph)
        self.time = [0] * len(self.solvers)      # next time index of each run
        self._deferred = None

    def time_stage(self, reps: int = 3) -> float:
        """Milliseconds per middle RK stage of the whole batch."""
        return _time_stage(self.h, reps)

    def info(self):
        ng = C.c_int()
        nb = C.c_long()
        check(lib().sparc_hip_batch_info(self.h, C.byref(ng), C.byref(nb)), "sparc_hip_batch_info")
        return ng.value, nb.value

    def advance(self, nsteps: int, defer_filter: bool = False, check_guard: bool = True):
        """Every run advances nsteps driver steps.  defer_filter leaves out
        the filters of the last step: the runs then hold the state the
        reference writes out (after STEP, before FILTER_VARS, driver.f90:
        198-200) until filter_deferred().  Returns {run index:
        DivergenceError} of the runs whose failure checks have tripped."""
        self.filter_deferred()
        n = len(self.solvers)
        if nsteps <= 0:
            return {}
        t0 = (C.c_int * n)(*self.time)
        amps = np.zeros((n, nsteps))
        for r, s in enumerate(self.solvers):
            if s.cfg.pulse:
                amps[r] = [s.amplitude(self.time[r] + q) for q in range(nsteps)]
        check(lib().sparc_hip_batch_advance(self.h, int(nsteps), t0, dptr(amps),
                                            DEFER_FILTER if defer_filter else 0),
              "sparc_hip_batch_advance")
        for r, s in enumerate(self.solvers):
            s._note(self.time[r], nsteps)
        self.time = [t + nsteps for t in self.time]
        if defer_filter:
            self._deferred = [t - 1 for t in self.time]
        return self.failures() if check_guard else {}

    def filter_deferred(self):
        """Apply the filters advance(defer_filter=True) left out."""
        if self._deferred is None:
            return
        t = (C.c_int * len(self._deferred))(*self._deferred)
        self._deferred = None
        check(lib().sparc_hip_batch_filter(self.h, t), "sparc_hip_batch_filter")

    def failures(self):
        """{run index: DivergenceError} of the runs whose failure checks
        have tripped (synchronises each run's context)."""
        out = {}
        for r, s in enumerate(self.solvers):
            failed, t, d, _ = s.guard()
            if failed:
                out[r] = DivergenceError(t, d)
        return out

    def close(self):
        h = getattr(self, "h", None)
        if h:
            lib().sparc_hip_batch_destroy(h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
