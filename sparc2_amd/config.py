"""Run configuration for the SPARC2 time-step path.

Field names follow the keys of the reference control file (``sparc_parameters``,
parsed by ``READ_SPARC_PARAMETERS``, initialize.f90:1385-1924).  Derived values
(cadences, step count, flags) follow the same routine.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional, Tuple

RSUN = 69598946770.0          # all_modules.f90:56
PI = 3.14159265358979         # all_modules.f90:56

FACES = ("xbot", "xtop", "ybot", "ytop", "zbot", "ztop")


@dataclasses.dataclass
class RunConfig:
    nx: int
    ny: int
    nz: int
    xlength: float                      # cm
    ylength: float                      # cm
    timestep: float = 1.0               # s
    solartime: float = 1.0              # h
    magnetic: int = 0                   # 0 quiet, 1 thermal, 2 MHD (MAGNETIC switch)
    # None: periodic in x and y unless USE_HPML is on
    periodic: Optional[bool] = None
    use_pml: bool = True
    use_hpml: bool = False
    stabilise: bool = True
    damping: bool = False               # DAMPING: FFT damping of the velocities (RHS.f90:166-172)
    # PML lines "N, RC, n, kappa, sp" per face, in FACES order
    npml: Tuple[int, int, int, int, int, int] = (0, 0, 0, 0, 10, 10)
    pmlc: Tuple[Tuple[float, float, float, float], ...] = ((9.2103, 2.0, 1.0, 0.1),) * 6
    filter_mode_xy: int = 1             # 1 -> 11-point (g = 5), 2 -> 7-point (g = 3)
    filter_mode_z: int = 1
    s_xy: float = 1.0
    s_z: float = 1.0
    t_xy: float = 30.0                  # s
    t_z: float = 10.0                   # s
    pulse: bool = True
    # PULSE = 0 driver (a forcing cube read from a FITS file, driver.f90:73-80,
    # 130-133): file, cadence (s), depth (cm), sigma (cm), amplitude (cm/s)
    forcingfunc: str = ""
    cadforcing: float = 60.0
    randomexcitedepth: float = -0.15e8
    randomexcitesigma: float = 0.1e8
    randomexciteampl: float = 1.0e4
    pulse_dir: int = 3
    pulse_t0: float = 1000.0 / 3.0      # s
    pulse_t1: float = 600.0
    pulse_st: float = 150.0
    pulse_px: float = 0.0               # cm
    pulse_py: float = 0.0
    pulse_pz: float = -0.15e8
    pulse_sr: float = 1.0e8
    pulse_sz: float = 1.0e8
    pulse_amp: float = 1.0e4
    sp_str: float = 0.0
    sp_length: float = 10.0e8
    sp_decay: float = 0.008
    maxva: float = 1.0e7                # cm/s (compared in cgs, initialize.f90:974)
    precision: str = "fp64"             # "fp64" or "fp32" (FP32 fields + Kahan RK update)

    def __post_init__(self):
        if self.periodic is None:
            self.periodic = not self.use_hpml
        elif self.periodic and self.use_hpml:
            # A control file with PERIODIC = 1 before USE_HPML = 1 leaves both
            # on in the reference (initialize.f90:1600-1624: the PERIODIC case
            # only checks a USE_HPML already read).  That combination (wrapped
            # x/y derivatives with x/y CPML faces) is not implemented here, so
            # it is refused rather than changed into another run.
            raise ValueError("PERIODIC and USE_HPML are both on (PERIODIC = 1 read before "
                             "USE_HPML = 1): periodic x/y boundaries with horizontal PMLs are "
                             "not supported; set PERIODIC = 0 or USE_HPML = 0")
        self.periodic = bool(self.periodic)
        npml = list(self.npml)
        if not self.use_hpml:
            npml[0:4] = [0, 0, 0, 0]    # initialize.f90:1626-1648
        if not self.use_pml:
            npml[4:6] = [0, 0]          # initialize.f90:1650-1660
        self.npml = tuple(int(v) for v in npml)
        self.pmlc = tuple(tuple(float(x) for x in row) for row in self.pmlc)

    # ---- derived values (initialize.f90:1548-1562, 1762-1768, 1894-1912) ----
    @property
    def g_xy_filt(self) -> int:
        return 5 if self.filter_mode_xy == 1 else 3

    @property
    def g_z_filt(self) -> int:
        return 5 if self.filter_mode_z == 1 else 3

    @property
    def xyfreq(self) -> int:
        t = max(self.t_xy, self.timestep)
        return int(math.floor(t / self.timestep + 1.0e-5))

    @property
    def zfreq(self) -> int:
        t = max(self.t_z, self.timestep)
        return int(math.floor(t / self.timestep + 1.0e-5))

    @property
    def maxtime(self) -> int:
        m = int(math.floor(self.solartime * 3600.0 / self.timestep))
        return 2 * (m // 2)

    @property
    def mhd(self) -> bool:
        return self.magnetic == 2

    @property
    def nv(self) -> int:
        return 8 if self.mhd else 5

    @property
    def npsi(self) -> int:
        return 6 if self.mhd else 2

    @property
    def npoints(self) -> int:
        return self.nx * self.ny * self.nz

    def face_shape(self, f: int) -> Tuple[int, int, int]:
        """(nz, ny, nx)-ordered shape of PML face ``f``'s arrays."""
        d = [self.nx, self.ny, self.nz]
        d[f // 2] = self.npml[f]
        return (d[2], d[1], d[0])

    def face_active(self, f: int) -> bool:
        on = self.use_hpml if f < 4 else self.use_pml
        return bool(on) and self.npml[f] > 0

    def validate(self) -> None:
        for n, name in ((self.nx, "nx"), (self.ny, "ny"), (self.nz, "nz")):
            if n < 16:
                raise ValueError(f"{name}={n}: need >= 16 cells per axis")
        for f in range(6):
            if self.face_active(f) and self.npml[f] < 2:
                raise ValueError(f"PML face {FACES[f]}: width must be >= 2")
        if self.filter_mode_xy not in (1, 2) or self.filter_mode_z not in (1, 2):
            raise ValueError("filter_mode must be 1 or 2")
        if self.precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        if self.pulse and self.pulse_dir not in (1, 2, 3):
            raise ValueError("pulse_dir must be 1, 2 or 3")

    def replace(self, **kw) -> "RunConfig":
        return dataclasses.replace(self, **kw)
