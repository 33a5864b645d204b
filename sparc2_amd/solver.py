"""Solver: one SPARC2 run resident on one GPU, driven through the C ABI.

Method names mirror the reference routines they replace:

    Solver.step(time)         STEP (step.f90:20-63), source at t = time*timestep
    Solver.filter_vars(time)  FILTER_VARS (derivatives.f90:659-703)
    Solver.ddx/ddy/ddz        ddx/ddy/ddz (derivatives.f90:10,226,451)
    Solver.filtx/filty/filtz  filtx/filty/filtz (derivatives.f90:119,345,565)
    Solver.norm_ke / diag     norm() + the negativity guards (driver.f90:161-183)
    Solver.run(n0, n1)        the driver loop body (driver.f90:123-200)

The reference's failure checks (driver.f90:175-196) run on the device in
every step's diagnostic reduction; ``advance`` reads the result once per
call (one synchronisation) and raises DivergenceError naming the failing
step.  In the default halt mode the device stops the run at that step, so
the state it holds is the one the reference dumps to full_XXXX.h5.

State arrays are numpy (nv, nz, ny, nx) FP64 == the reference's Fortran
``a(nx, ny, nz, nv)``.  PML bases are per face, (npsi, *face_shape).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import SparcBg, SparcCfg, check, dptr, lib
from .config import RunConfig
from .forcing import ForcingSeries, source_dep
from .initialise import (SetupState, damping_rates, initialise_rhs, layout_for_device, prepare,
                         source_amplitude, source_profiles)


# Autotuned tile per grid signature (see Solver.autotune).
_TILE_CACHE: Dict[tuple, tuple] = {}


class DivergenceError(RuntimeError):
    """The reference's failure checks tripped (driver.f90:175-196): norm_KE
    above 1e3, or a non-positive rho+rho0 or p+p0, at time index ``time``;
    ``diag`` = (norm_KE, min rho+rho0, min p+p0) of that step, ``dump`` the
    full_XXXX.h5 written, if any."""

    def __init__(self, time: int, diag, dump: Optional[str] = None):
        self.time, self.diag, self.dump = time, tuple(float(v) for v in diag), dump
        msg = (f"step {time}: norm_KE={self.diag[0]:.5e} min(rho+rho0)={self.diag[1]:.5e} "
               f"min(p+p0)={self.diag[2]:.5e}")
        if dump:
            msg += f"; state written to {dump}"
        super().__init__(msg)


class Solver:
    def __init__(self, cfg: RunConfig, raw_bg: Dict[str, np.ndarray], device: int = 0,
                 tile: Sequence[int] = (0, 0, 0), setup: Optional[SetupState] = None):
        self.cfg = cfg
        self.st = setup if setup is not None else prepare(cfg, raw_bg)
        st = self.st
        L = lib()
        self._keep = []
        c = SparcCfg()
        c.nx, c.ny, c.nz = cfg.nx, cfg.ny, cfg.nz
        c.option = 3 if cfg.mhd else 1
        c.precision = _lib.PREC_FP32 if cfg.precision == "fp32" else _lib.PREC_FP64
        c.periodic = int(cfg.periodic)
        c.use_pml = int(cfg.use_pml)
        c.use_hpml = int(cfg.use_hpml)
        for f in range(6):
            c.npml[f] = cfg.npml[f]
        c.g_xy_filt, c.g_z_filt = cfg.g_xy_filt, cfg.g_z_filt
        c.s_xy, c.s_z = cfg.s_xy, cfg.s_z
        c.xyfreq, c.zfreq = cfg.xyfreq, cfg.zfreq
        c.pulse_dir = cfg.pulse_dir if cfg.pulse else 0
        c.bg3d = int(st.bg3d)
        c.deltat = st.deltat
        for s in range(5):
            c.betas[s] = float(st.betas[s])
        c.stretchx, c.stretchy = st.stretchx, st.stretchy
        self._stretch = np.ascontiguousarray(st.stretch, dtype=np.float64)
        c.stretch = dptr(self._stretch)
        c.device = device
        for q in range(3):
            c.tile[q] = int(tile[q])
        self._cfg = c
        self._lanes_env = os.environ.get("SPARC_LANE_POINTS", "")
        h = C.c_void_p()
        check(L.sparc_hip_create(C.byref(c), C.byref(h)), "sparc_hip_create")
        self.h = h
        self.nv = cfg.nv
        self.npsi = cfg.npsi
        self.shape = (cfg.nz, cfg.ny, cfg.nx)
        if setup is None:
            initialise_rhs(st, self)
        self._upload_background()
        self.put_state(np.zeros((self.nv,) + self.shape))
        # PULSE = 0: the forcing series of the FITS cube (use_forcing attaches
        # another one); every step uploads its forcing map first
        self.series: Optional[ForcingSeries] = None
        self._sdep_sent = False
        # steps issued, and which time index each one had: (first count,
        # first time, n) per contiguous stretch -- maps the guard's step
        # count back to a time index
        self._count = 0
        self._segs = []
        if not cfg.pulse and cfg.forcingfunc:
            self.use_forcing(ForcingSeries.from_file(st))

    # ---------------------------------------------------------------- set-up
    def _upload_background(self):
        st = self.st
        dev = layout_for_device(st)
        b = SparcBg()
        for name, v in dev.items():
            setattr(b, name, dptr(v))
            self._keep.append(v)
        for f in range(6):
            if f in st.pml:
                for key, fld in (("a", "pml_a"), ("b", "pml_b"), ("kappa", "pml_kappa")):
                    v = np.ascontiguousarray(st.pml[f][key], dtype=np.float64)
                    self._keep.append(v)
                    getattr(b, fld)[f] = dptr(v)
        if self.cfg.pulse:
            sx, sy, sz = (np.ascontiguousarray(v) for v in source_profiles(st))
            self._keep += [sx, sy, sz]
            b.src_x, b.src_y, b.src_z = dptr(sx), dptr(sy), dptr(sz)
        b.dimc, b.dimrho = st.dimc, st.dimrho
        check(lib().sparc_hip_set_background(self.h, C.byref(b)), "sparc_hip_set_background")
        if self.cfg.damping:
            # DAMPING: INIT_DAMPING's rates, C (ny, nx/2+1) = Fortran (nx/2+1, ny)
            self._damp_rates = np.ascontiguousarray(damping_rates(st), dtype=np.float64)
            check(lib().sparc_hip_set_damping(self.h, dptr(self._damp_rates)), "sparc_hip_set_damping")

    def close(self):
        h = getattr(self, "h", None)
        if h:
            lib().sparc_hip_destroy(h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ----------------------------------------------------------------- state
    def face_shape(self, f):
        return (self.npsi,) + self.cfg.face_shape(f)

    def put_state(self, a, psi: Optional[Dict[int, np.ndarray]] = None):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != (self.nv,) + self.shape:
            raise ValueError(f"state shape {a.shape} != {(self.nv,) + self.shape}")
        ps = (_lib.DP * 6)()
        keep = [a]
        for f in range(6):
            if psi is not None and f in psi and self.cfg.face_active(f):
                v = np.ascontiguousarray(psi[f], dtype=np.float64)
                if v.size != int(np.prod(self.face_shape(f))):
                    raise ValueError(f"psi face {f} has {v.size} values")
                keep.append(v)
                ps[f] = dptr(v)
        check(lib().sparc_hip_put_state(self.h, dptr(a), ps), "sparc_hip_put_state")

    def get_state(self, with_psi: bool = False):
        a = np.zeros((self.nv,) + self.shape)
        ps = (_lib.DP * 6)()
        out = {}
        if with_psi:
            for f in range(6):
                if self.cfg.face_active(f):
                    out[f] = np.zeros(self.face_shape(f))
                    ps[f] = dptr(out[f])
        check(lib().sparc_hip_get_state(self.h, dptr(a), ps), "sparc_hip_get_state")
        return (a, out) if with_psi else a

    # ------------------------------------------------------------- time step
    def amplitude(self, time: int) -> float:
        return source_amplitude(self.st, time)

    def use_forcing(self, series: Optional[ForcingSeries]):
        """Drive the run with a PULSE = 0 forcing series (None: back to the
        pulse settings of the run config)."""
        self.series = series
        self._sdep_sent = False
        if series is None:
            check(lib().sparc_hip_set_forcing(self.h, None, None), "sparc_hip_set_forcing")

    def _push_forcing(self, time: int):
        fmap = np.ascontiguousarray(self.series.forcing(time), dtype=np.float64)
        sdep = None
        if not self._sdep_sent:
            self._sdep = np.ascontiguousarray(source_dep(self.st), dtype=np.float64)
            sdep = dptr(self._sdep)
        check(lib().sparc_hip_set_forcing(self.h, dptr(fmap), sdep), "sparc_hip_set_forcing")
        self._sdep_sent = True

    def _note(self, time0: int, n: int):
        if n <= 0:
            return
        if self._segs:
            c0, t0, m = self._segs[-1]
            if c0 + m == self._count + 1 and t0 + m == time0:
                self._segs[-1] = (c0, t0, m + n)
                self._count += n
                return
        self._segs.append((self._count + 1, int(time0), n))
        self._count += n

    def time_of_step(self, count: int) -> int:
        """Time index of the count-th step this context took (from 1)."""
        for c0, t0, m in reversed(self._segs):
            if c0 <= count < c0 + m:
                return t0 + count - c0
        raise KeyError(count)

    def step(self, time: int):
        """STEP at time index ``time`` (asynchronous; no guard check)."""
        if self.series is not None:
            self._push_forcing(time)
        self._note(time, 1)
        check(lib().sparc_hip_step(self.h, self.amplitude(time)), "sparc_hip_step")

    def filter_vars(self, time: int):
        check(lib().sparc_hip_filter(self.h, int(time)), "sparc_hip_filter")

    def advance(self, n0: int, n1: int, check_guard: bool = True, filter_last: bool = True):
        """The driver loop body for nn = n0..n1 (source, STEP, FILTER_VARS
        when nn != 0).  filter_last=False leaves out the filter of step n1
        (the caller writes the reference's outputs, then calls
        filter_vars(n1)).  With check_guard the failure checks of every step
        are read at the end (one synchronisation) and a failure raises
        DivergenceError."""
        if n1 < n0:
            return
        last = n1 if filter_last else n1 - 1
        if self.series is not None:
            # the forcing map changes every step: the driver loop runs here
            for t in range(n0, n1 + 1):
                self.step(t)
                if t != 0 and t <= last:
                    self.filter_vars(t)
        else:
            if last >= n0:
                amps = np.array([self.amplitude(t) for t in range(n0, last + 1)], dtype=np.float64)
                self._note(n0, last - n0 + 1)
                check(lib().sparc_hip_advance(self.h, int(n0), int(last), dptr(amps)),
                      "sparc_hip_advance")
            if last < n1:
                self.step(n1)
        if check_guard:
            self.check_guard()

    # ----------------------------------------------------------- guards
    def set_guard(self, mode: int = _lib.GUARD_HALT):
        """SPARC_GUARD_HALT (stop at the failing step, the default),
        _RECORD (note the first failure, keep stepping) or _OFF; clears a
        recorded failure."""
        check(lib().sparc_hip_set_guard(self.h, int(mode)), "sparc_hip_set_guard")

    def guard(self):
        """(failed, time index of the failing step or None, its diag, steps
        completed on the device).  Synchronises."""
        tripped = C.c_int()
        at = C.c_long()
        done = C.c_long()
        d = np.zeros(3)
        check(lib().sparc_hip_guard(self.h, C.byref(tripped), C.byref(at), C.byref(done), dptr(d)),
              "sparc_hip_guard")
        if not tripped.value:
            return False, None, None, done.value
        return True, self.time_of_step(at.value), d, done.value

    def check_guard(self, dump_dir: Optional[str] = None, simname: str = ""):
        """Raise DivergenceError if a step failed the reference's checks;
        with dump_dir the state the run stopped at is first written there as
        <simname>full_XXXX.h5 (WRITE_OUT_FULL_STATE with fail set,
        all_modules.f90:1090-1100)."""
        failed, t, d, _ = self.guard()
        if not failed:
            return
        path = None
        if dump_dir is not None:
            from . import h5io
            path = os.path.join(dump_dir, h5io.file_name(simname, "full", None))
            h5io.write_full(path, self, time=t)
        raise DivergenceError(t, d, path)

    def diag(self) -> np.ndarray:
        out = np.zeros(3)
        check(lib().sparc_hip_diag(self.h, dptr(out)), "sparc_hip_diag")
        return out

    def run(self, n0: int, n1: int, check_every: int = 1,
            dump_dir: Optional[str] = None, simname: str = "") -> np.ndarray:
        """Driver loop nn = n0..n1 returning the diagnostics (norm_KE,
        min rho+rho0, min p+p0) of every ``check_every``-th step (a
        synchronisation each); the failure checks run on the device every
        step and a failure raises DivergenceError (driver.f90:175-196), after
        writing the dump when ``dump_dir`` is given."""
        hist = []
        for nn in range(n0, n1 + 1):
            self.step(nn)
            if (nn - n0) % check_every == 0 or nn == n1:
                self.check_guard(dump_dir, simname)
                hist.append(self.diag())
            if nn != 0:
                self.filter_vars(nn)
        self.check_guard(dump_dir, simname)
        return np.array(hist)

    def sync(self):
        check(lib().sparc_hip_sync(self.h), "sparc_hip_sync")

    # ------------------------------------------------------------- operators
    def _op(self, fn, x, *extra):
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.size != int(np.prod(self.shape)):
            raise ValueError("operator input must hold nx*ny*nz values")
        out = np.zeros(self.shape)
        check(fn(self.h, dptr(x), dptr(out), *extra), fn.__name__)
        return out

    def ddx(self, x, bc):
        return self._op(lib().sparc_hip_ddx, x, int(bc))

    def ddy(self, x, bc):
        return self._op(lib().sparc_hip_ddy, x, int(bc))

    def ddz(self, x, bc=1):
        return self._op(lib().sparc_hip_ddz, x, int(bc))

    def filtx(self, x):
        return self._op(lib().sparc_hip_filtx, x)

    def filty(self, x):
        return self._op(lib().sparc_hip_filty, x)

    def filtz(self, x):
        return self._op(lib().sparc_hip_filtz, x)

    def alfven_search(self, red, b2, rho0, dimb, dimrho, maxva) -> float:
        """Search temperature of the Alfven limiter, maximum over the grid on
        the device (initialize.f90:962-986)."""
        arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (red, b2, rho0)]
        t = C.c_double()
        check(lib().sparc_hip_alfven_search(self.h, *(dptr(v) for v in arrs), float(dimb),
                                            float(dimrho), float(maxva), C.byref(t)),
              "sparc_hip_alfven_search")
        return t.value

    def get_slab(self, field: int, k0: int, k1: int) -> np.ndarray:
        """Planes k0..k1 of one state field (the chunk output of WRITE_OUT_CHUNK)."""
        out = np.zeros((k1 - k0 + 1, self.cfg.ny, self.cfg.nx))
        check(lib().sparc_hip_get_slab(self.h, int(field), int(k0), int(k1), dptr(out)),
              "sparc_hip_get_slab")
        return out

    # ----------------------------------------------------------- measurement
    def timing(self, enable: bool):
        check(lib().sparc_hip_timing(self.h, int(enable)), "sparc_hip_timing")

    def kernel_time(self, kind: int):
        n = C.c_long()
        ms = C.c_double()
        check(lib().sparc_hip_kernel_time(self.h, kind, C.byref(n), C.byref(ms)), "kernel_time")
        return n.value, ms.value

    def tile(self):
        t = (C.c_int * 3)()
        nb = C.c_long()
        check(lib().sparc_hip_tile(self.h, t, C.byref(nb)), "sparc_hip_tile")
        return tuple(t), nb.value

    def set_tile(self, tile: Sequence[int]):
        t = (C.c_int * 3)(*[int(v) for v in tile])
        check(lib().sparc_hip_set_tile(self.h, t), "sparc_hip_set_tile")

    def tile_key(self) -> tuple:
        """What the best tile depends on: grid, physics variant, precision,
        boundary kind, CPML widths and the points per lane (SPARC_LANE_POINTS
        read when the context was made)."""
        c = self.cfg
        return (c.nx, c.ny, c.nz, c.magnetic, c.precision, bool(c.periodic), bool(c.use_hpml),
                tuple(c.npml), bool(self.st.bg3d), self._lanes_env)

    def autotune(self, reps: int = 3, cache: bool = True):
        """Pick the fastest tile shape for this grid by timing a middle RK
        stage per candidate (the run's state is unchanged); runs with the same
        tile_key reuse the answer.  Returns (tile, ms per stage or None)."""
        key = self.tile_key()
        if cache and key in _TILE_CACHE:
            self.set_tile(_TILE_CACHE[key])
            return _TILE_CACHE[key], None
        t = (C.c_int * 3)()
        ms = C.c_double()
        check(lib().sparc_hip_autotune(self.h, int(reps), t, C.byref(ms)), "sparc_hip_autotune")
        _TILE_CACHE[key] = tuple(t)
        return tuple(t), ms.value
