"""Batched form: many independent runs advanced by shared persistent launches.

    b = Batch([Solver(cfg_i, bg_i) for ...], use_graph=True)
    b.advance(1000)          # every run advances 1000 driver steps

Each run keeps its own time index; its source amplitudes (driver.f90:140-142)
and filter cadence (derivatives.f90:669,690) follow that index.  The runs'
states stay device-resident in their own contexts; ``Solver.get_state`` and
``Solver.diag`` read them as usual.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence

import numpy as np

from ._lib import check, dptr, lib
from .solver import Solver


class Batch:
    def __init__(self, solvers: Sequence[Solver], use_graph: bool = True, autotune: bool = False):
        self.solvers: List[Solver] = list(solvers)
        if not self.solvers:
            raise ValueError("empty batch")
        if autotune:
            # one timing sweep per distinct grid; runs alike reuse it
            for s in self.solvers:
                s.autotune()
        n = len(self.solvers)
        hs = (C.c_void_p * n)(*[s.h.value for s in self.solvers])
        h = C.c_void_p()
        check(lib().sparc_hip_batch_create(hs, n, int(use_graph), C.byref(h)), "sparc_hip_batch_create")
        self.h = h
        self.time = [0] * n      # next time index of each run

    def info(self):
        ng = C.c_int()
        nb = C.c_long()
        check(lib().sparc_hip_batch_info(self.h, C.byref(ng), C.byref(nb)), "sparc_hip_batch_info")
        return ng.value, nb.value

    def advance(self, nsteps: int):
        n = len(self.solvers)
        t0 = (C.c_int * n)(*self.time)
        amps = np.zeros((n, nsteps))
        for r, s in enumerate(self.solvers):
            if s.cfg.pulse:
                amps[r] = [s.amplitude(self.time[r] + q) for q in range(nsteps)]
        check(lib().sparc_hip_batch_advance(self.h, int(nsteps), t0, dptr(amps)),
              "sparc_hip_batch_advance")
        self.time = [t + nsteps for t in self.time]

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
