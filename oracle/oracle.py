"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

Arrays use numpy shape (nz, ny, nx) C order == the reference's Fortran
(nx, ny, nz) layout; multi-field state is (nv, nz, ny, nx).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SPARC_ORACLE_LIB names another build of the oracle: the mutation tests load
# copies with planted mistakes through it.
LIB_PATH = os.environ.get("SPARC_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")


class OrcCfg(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("magnetic", C.c_int),
        ("periodic", C.c_int), ("use_pml", C.c_int), ("use_hpml", C.c_int),
        ("stabilise", C.c_int),
        ("npml", C.c_int * 6),
        ("pmlc", (C.c_double * 4) * 6),
        ("xlength", C.c_double), ("ylength", C.c_double), ("timestep", C.c_double),
        ("g_xy_filt", C.c_int), ("g_z_filt", C.c_int),
        ("s_xy", C.c_double), ("s_z", C.c_double), ("t_xy", C.c_double), ("t_z", C.c_double),
        ("pulse_dir", C.c_int),
        ("pulse_t0", C.c_double), ("pulse_t1", C.c_double), ("pulse_st", C.c_double),
        ("pulse_px", C.c_double), ("pulse_py", C.c_double), ("pulse_pz", C.c_double),
        ("pulse_sr", C.c_double), ("pulse_sz", C.c_double), ("pulse_amp", C.c_double),
        ("sp_str", C.c_double), ("sp_length", C.c_double), ("sp_decay", C.c_double),
        ("maxva", C.c_double),
    ]


def build(force: bool = False) -> str:
    if os.environ.get("SPARC_ORACLE_LIB"):
        return LIB_PATH
    src = os.path.join(HERE, "sparc_oracle.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "liboracle.so"],
                       check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(C.c_double)
        vp = C.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(OrcCfg)] + [P] * 8
        L.orc_destroy.argtypes = [vp]
        L.orc_last_error.restype = C.c_char_p
        for n in ("orc_nv", "orc_npsi"):
            getattr(L, n).argtypes = [vp]
            getattr(L, n).restype = C.c_int
        L.orc_face_size.argtypes = [vp, C.c_int]
        L.orc_face_size.restype = C.c_long
        L.orc_scalar.argtypes = [vp, C.c_char_p]
        L.orc_scalar.restype = C.c_double
        L.orc_array_len.argtypes = [vp, C.c_char_p]
        L.orc_array_len.restype = C.c_long
        L.orc_get_array.argtypes = [vp, C.c_char_p, P]
        for n in ("orc_put_state", "orc_get_state", "orc_get_temp_step", "orc_get_scr"):
            getattr(L, n).argtypes = [vp, P]
        L.orc_put_psi.argtypes = [vp, C.c_int, P]
        L.orc_get_psi.argtypes = [vp, C.c_int, P]
        L.orc_forcing.argtypes = [vp, C.c_int]
        L.orc_step.argtypes = [vp]
        L.orc_filter_vars.argtypes = [vp, C.c_int]
        L.orc_diag.argtypes = [vp, P]
        L.orc_run.argtypes = [vp, C.c_int, C.c_int, P]
        L.orc_rhs_stage.argtypes = [vp, C.c_int]
        L.orc_set_damping.argtypes = [vp, C.c_int]
        L.orc_set_damping.restype = C.c_int
        L.orc_damp_term.argtypes = [vp, P, P]
        L.orc_set_forcing_map.argtypes = [vp, P, P]
        L.orc_set_forcing_map.restype = C.c_int
        for n in ("orc_ddx", "orc_ddy", "orc_ddz"):
            getattr(L, n).argtypes = [vp, P, P, C.c_int]
            getattr(L, n).restype = C.c_int
        for n in ("orc_filtx", "orc_filty", "orc_filtz"):
            getattr(L, n).argtypes = [vp, P, P]
        L.orc_coeffs.argtypes = [P, P, P]
        L.orc_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def coeffs():
    dp = np.zeros(66)
    f11 = np.zeros(11)
    f7 = np.zeros(7)
    lib().orc_coeffs(_p(dp), _p(f11), _p(f7))
    return dp.reshape(6, 11), f11, f7


def make_cfg(cfg) -> OrcCfg:
    """Build the C struct from any object with RunConfig-like attributes."""
    o = OrcCfg()
    o.nx, o.ny, o.nz = cfg.nx, cfg.ny, cfg.nz
    o.magnetic = int(cfg.magnetic)
    o.periodic = int(bool(cfg.periodic))
    o.use_pml = int(bool(cfg.use_pml))
    o.use_hpml = int(bool(cfg.use_hpml))
    o.stabilise = int(bool(cfg.stabilise))
    for f in range(6):
        o.npml[f] = int(cfg.npml[f])
        for q in range(4):
            o.pmlc[f][q] = float(cfg.pmlc[f][q])
    o.xlength, o.ylength, o.timestep = cfg.xlength, cfg.ylength, cfg.timestep
    o.g_xy_filt = 5 if cfg.filter_mode_xy == 1 else 3
    o.g_z_filt = 5 if cfg.filter_mode_z == 1 else 3
    o.s_xy, o.s_z, o.t_xy, o.t_z = cfg.s_xy, cfg.s_z, cfg.t_xy, cfg.t_z
    o.pulse_dir = int(cfg.pulse_dir) if cfg.pulse else 0
    for n in ("pulse_t0", "pulse_t1", "pulse_st", "pulse_px", "pulse_py", "pulse_pz",
              "pulse_sr", "pulse_sz", "pulse_amp", "sp_str", "sp_length", "sp_decay", "maxva"):
        setattr(o, n, float(getattr(cfg, n)))
    return o


class Oracle:
    """One reference run on the CPU (the whole domain, virtual multi-block edges)."""

    def __init__(self, cfg, bg):
        self.cfg = cfg
        L = lib()
        self._c = make_cfg(cfg)
        self._keep = {}

        def arr(name):
            v = bg.get(name)
            if v is None:
                return None
            v = np.ascontiguousarray(v, dtype=np.float64)
            self._keep[name] = v
            return _p(v)

        self.h = L.orc_create(C.byref(self._c), arr("rho0"), arr("pre0"), arr("cs0"),
                              arr("bx0"), arr("by0"), arr("bz0"), arr("zaxis"), arr("grav"))
        if not self.h:
            raise ValueError(L.orc_last_error().decode())
        self.nv = L.orc_nv(self.h)
        self.npsi = L.orc_npsi(self.h)
        self.shape = (cfg.nz, cfg.ny, cfg.nx)
        if getattr(cfg, "damping", False):
            self.set_damping(True)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().orc_destroy(h)
            self.h = None

    # -- DAMPING --
    def set_damping(self, on: bool):
        if lib().orc_set_damping(self.h, int(on)) != 0:
            raise MemoryError("damping rates")

    def damp_term(self, v):
        """damp_velocity of one (nz, ny, nx) field with the current rates."""
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros_like(v)
        lib().orc_damp_term(self.h, _p(v), _p(out))
        return out

    # -- PULSE = 0 --
    def set_forcing_map(self, fmap, sdep=None):
        """This step's forcing (ny, nx) and source_dep (nz); None: the pulse."""
        if fmap is None:
            lib().orc_set_forcing_map(self.h, None, None)
            return
        self._fmap = np.ascontiguousarray(fmap, dtype=np.float64)
        self._sdep = None if sdep is None else np.ascontiguousarray(sdep, dtype=np.float64)
        if lib().orc_set_forcing_map(self.h, _p(self._fmap),
                                     None if self._sdep is None else _p(self._sdep)) != 0:
            raise MemoryError("forcing map")

    # -- set-up state --
    def scalar(self, name):
        return lib().orc_scalar(self.h, name.encode())

    def array(self, name):
        n = lib().orc_array_len(self.h, name.encode())
        if n < 0:
            raise KeyError(name)
        out = np.zeros(n)
        lib().orc_get_array(self.h, name.encode(), _p(out))
        if n == self.cfg.nx * self.cfg.ny * self.cfg.nz:
            return out.reshape(self.shape)
        return out

    def face_size(self, f):
        return lib().orc_face_size(self.h, f)

    # -- state --
    def put_state(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.nv * np.prod(self.shape)
        lib().orc_put_state(self.h, _p(a))

    def get_state(self):
        out = np.zeros((self.nv,) + self.shape)
        lib().orc_get_state(self.h, _p(out))
        return out

    def temp_step(self):
        out = np.zeros((self.nv,) + self.shape)
        lib().orc_get_temp_step(self.h, _p(out))
        return out

    def scr(self):
        out = np.zeros((self.nv,) + self.shape)
        lib().orc_get_scr(self.h, _p(out))
        return out

    def put_psi(self, f, psi):
        psi = np.ascontiguousarray(psi, dtype=np.float64)
        assert psi.size == self.npsi * self.face_size(f)
        lib().orc_put_psi(self.h, f, _p(psi))

    def get_psi(self, f):
        out = np.zeros(self.npsi * self.face_size(f))
        lib().orc_get_psi(self.h, f, _p(out))
        return out

    # -- driver pieces --
    def forcing(self, time):
        lib().orc_forcing(self.h, int(time))

    def step(self):
        lib().orc_step(self.h)

    def rhs_stage(self, stage):
        lib().orc_rhs_stage(self.h, int(stage))

    def filter_vars(self, time):
        lib().orc_filter_vars(self.h, int(time))

    def diag(self):
        out = np.zeros(3)
        lib().orc_diag(self.h, _p(out))
        return out

    def run(self, n0, n1):
        hist = np.zeros((n1 - n0 + 1, 3))
        lib().orc_run(self.h, int(n0), int(n1), _p(hist))
        return hist

    # -- operators --
    def _op(self, name, x, *extra):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros_like(x)
        rc = getattr(lib(), name)(self.h, _p(x), _p(out), *extra)
        if rc not in (None, 0):
            raise ValueError(f"{name}: bad boundary code")
        return out

    def ddx(self, x, bc):
        return self._op("orc_ddx", x, int(bc))

    def ddy(self, x, bc):
        return self._op("orc_ddy", x, int(bc))

    def ddz(self, x, bc=1):
        return self._op("orc_ddz", x, int(bc))

    def filtx(self, x):
        return self._op("orc_filtx", x)

    def filty(self, x):
        return self._op("orc_filty", x)

    def filtz(self, x):
        return self._op("orc_filtz", x)
