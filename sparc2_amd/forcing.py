"""The PULSE = 0 driver: a forcing function read from a FITS cube, interpolated
in time, added to the vz right-hand side with a depth profile.

* ``read_fits_cube`` / ``write_fits_cube`` -- readfits_xy_distribute
  (all_modules.f90:470-517): the cube vr(nx, ny, num_steps), via
  libsparc_fits.so (CFITSIO).
* ``num_steps``   -- the cube depth the driver reads (driver.f90:38-39).
* ``source_dep``  -- the depth profile (all_modules.f90:289-301): a Gaussian
  around the grid plane nearest 1 + randomexcitedepth/Rsun.
* ``ForcingSeries`` -- vr scaled by randomexciteampl/dimc (driver.f90:79) and
  LAGRANGE_INTERP (all_modules.f90:1892-1950): 7-point Lagrange
  interpolation at t = time*timestep over the cube planes at the cadence
  cadforcing, times source_damp (driver.f90:131-132).

The time step then adds forcing(x, y) * source_dep(z) to RHSv_z
(RHS.f90:146-147, 640-641) in place of the pulse.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from typing import Optional

import numpy as np

from .config import RSUN

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libsparc_fits.so")
_lib = None


class FitsError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise FitsError(f"{LIB} is not built (python -m sparc2_amd.build; needs CFITSIO)")
        L = C.CDLL(LIB)
        DP = C.POINTER(C.c_double)
        L.sparc_fits_read_cube.argtypes = [C.c_char_p, C.c_long, C.c_long, C.c_long, DP]
        L.sparc_fits_write_cube.argtypes = [C.c_char_p, C.c_long, C.c_long, C.c_long, DP]
        L.sparc_fits_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def read_fits_cube(path: str, nx: int, ny: int, nsteps: int) -> np.ndarray:
    """vr as a C-order array (nsteps, ny, nx) = Fortran vr(nx, ny, nsteps)."""
    out = np.empty((nsteps, ny, nx))
    rc = lib().sparc_fits_read_cube(path.encode(), nx, ny, nsteps, out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise FitsError(lib().sparc_fits_last_error().decode())
    return out


def write_fits_cube(path: str, cube: np.ndarray):
    c = np.ascontiguousarray(cube, dtype=np.float64)
    nsteps, ny, nx = c.shape
    rc = lib().sparc_fits_write_cube(path.encode(), nx, ny, nsteps, c.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise FitsError(lib().sparc_fits_last_error().decode())


def num_steps(cfg) -> int:
    """driver.f90:38-39."""
    n = math.floor(float(cfg.maxtime) * cfg.timestep / cfg.cadforcing) + 2
    return 2 * (n // 2) + 2


def source_dep(st) -> np.ndarray:
    """all_modules.f90:289-301: the plane e_rad nearest 1 + depth/Rsun (the
    last of equally near ones), then exp(-0.5 (z - z_e)^2 / 2 / (sigma/Rsun)^2)."""
    cfg, z = st.cfg, st.zaxis
    temp2 = 1.0 + cfg.randomexcitedepth / RSUN
    temp = 1.0
    e_rad = 0
    for kk in range(cfg.nz):
        if abs(z[kk] - temp2) <= temp:
            temp = abs(z[kk] - temp2)
            e_rad = kk
    return np.exp(-0.5 * (z - z[e_rad]) ** 2 / 2.0 / (cfg.randomexcitesigma / RSUN) ** 2)


class ForcingSeries:
    """The forcing of each driver step: LAGRANGE_INTERP over the scaled cube,
    times source_damp.  Like the reference it keeps the seven weighted planes
    (LC0..LC6) of the current cadence interval and recomputes them only when
    the interval advances."""

    def __init__(self, st, cube: np.ndarray):
        cfg = st.cfg
        if cube.shape[1:] != (cfg.ny, cfg.nx):
            raise ValueError(f"forcing cube planes {cube.shape[1:]} != {(cfg.ny, cfg.nx)}")
        self.cfg = cfg
        self.vr = (cfg.randomexciteampl / st.dimc) * np.asarray(cube, dtype=np.float64)
        self.nsteps = cube.shape[0]
        self.damp = st.source_damp if st.source_damp is not None else np.ones((cfg.ny, cfg.nx))
        self.time_old = 0
        self.lc = None
        self.x = None

    @classmethod
    def from_file(cls, st):
        cfg = st.cfg
        return cls(st, read_fits_cube(cfg.forcingfunc, cfg.nx, cfg.ny, num_steps(cfg)))

    def _plane(self, p: int) -> np.ndarray:
        """vr(:, :, p), 1-based as in the reference."""
        if not 1 <= p <= self.nsteps:
            raise IndexError(f"forcing plane {p} outside the cube's {self.nsteps}")
        return self.vr[p - 1]

    def lagrange(self, time: int) -> np.ndarray:
        """LAGRANGE_INTERP (all_modules.f90:1892-1950) at the step `time`."""
        cad, dt = self.cfg.cadforcing, self.cfg.timestep
        xt = float(time) * dt / cad
        time1 = math.floor(xt) + 1
        xt = time * dt
        if self.time_old < time1:
            x = [(time1 - 4 + q) * cad for q in range(7)]
            lc = [0.0] * 7
            zero = np.zeros_like(self.vr[0])

            def den(q):
                d = 1.0
                for p in range(7):
                    if p != q:
                        d = d * (x[q] - x[p])
                return d
            lc[0] = self._plane(time1 - 3) / den(0) if x[0] >= 0 else zero
            lc[1] = self._plane(time1 - 2) / den(1) if x[1] >= 0 else zero
            lc[2] = self._plane(time1 - 1) / den(2) if x[2] >= 0 else zero
            lc[3] = self._plane(time1) / den(3) if x[3] >= 0 else zero
            lc[4] = self._plane(time1 + 1) / den(4)
            lc[5] = self._plane(time1 + 2) / den(5) if time1 < self.nsteps else zero
            lc[6] = self._plane(time1 + 3) / den(6) if time1 < self.nsteps - 1 else zero
            self.lc, self.x = lc, x
            self.time_old = time1
        out = 0.0
        for q in range(7):
            term = self.lc[q]
            for p in range(7):
                if p != q:
                    term = term * (xt - self.x[p])
            out = out + term
        return out

    def forcing(self, time: int) -> np.ndarray:
        """forcing = LAGRANGE_INTERP * source_damp (driver.f90:131-132)."""
        return self.lagrange(time) * self.damp
