"""The reference's HDF5 files from the Python host, through libsparc_h5.so.

* ``read_bkg`` / ``write_bkg``: the background input bkg.h5 that
  READ_IN_HDF5 reads (all_modules.f90:179-282): rho0, pre0, cs0, bx0/by0/bz0
  (magnetic runs), zaxis (Mm) and grav (nz values), root attribute xyzdims.
  ``write_bkg`` gives a generated background the layout the reference reads.
* ``write_full`` / ``read_full``: the snapshot and restart full_NNNN.h5
  (WRITE_OUT_FULL_STATE, all_modules.f90:1072-1423; READ_IN_FULL_STATE,
  741-1069): pre, rho, vx, vy, vz (+ bx, by, bz) in physical units and the
  CPML bases PML_ZTOP, PML_ZBOT, PML_XTOP, PML_XBOT, PML_YTOP, PML_YBOT, root
  attributes xyzdims and time (hours).
* ``write_chunk``: the chunk output chunk_NNNN.h5 (WRITE_OUT_CHUNK,
  all_modules.f90:1566-1799): xaxis, yaxis (cm), zaxis (Mm) and vx, vy, vz,
  rho, pre (+ bx, by, bz) over a sub-box, root attributes xyzdims and time.

Values are scaled with the reference's expressions, in its order of
operations (``a*dimrho*dimc**2`` on writing, ``/dimrho/dimc**2`` on
reading).  The displacement formulation (DISPL) is outside this path.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional, Sequence

import numpy as np

from .config import RSUN

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libsparc_h5.so")

# face index -> dataset name (faces: xbot, xtop, ybot, ytop, zbot, ztop)
PML_NAMES = {0: "PML_XBOT", 1: "PML_XTOP", 2: "PML_YBOT", 3: "PML_YTOP", 4: "PML_ZBOT", 5: "PML_ZTOP"}

_lib = None


class H5Error(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise H5Error(f"{LIB} is not built (python -m sparc2_amd.build; needs HDF5)")
        L = C.CDLL(LIB)
        LP = C.POINTER(C.c_long)
        DP = C.POINTER(C.c_double)
        IP = C.POINTER(C.c_int)
        L.sparc_h5_create.argtypes = [C.c_char_p, IP, C.c_double]
        L.sparc_h5_attrs.argtypes = [C.c_char_p, IP, DP]
        L.sparc_h5_has.argtypes = [C.c_char_p, C.c_char_p]
        L.sparc_h5_write.argtypes = [C.c_char_p, C.c_char_p, C.c_int, LP, DP, C.c_double]
        L.sparc_h5_read.argtypes = [C.c_char_p, C.c_char_p, C.c_int, LP, DP, C.c_double]
        L.sparc_h5_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _check(rc, what):
    if rc < 0:
        raise H5Error(f"{what}: {lib().sparc_h5_last_error().decode()}")
    return rc


def file_name(simname: str, kind: str, n: Optional[int], width: int = 4) -> str:
    """``simname + kind + '_' + NNNN + '.h5'``: n zero-padded to `width`
    digits (convert_to_string, all_modules.f90:710-737, with the reference's
    timestamp_size), or 'XXXX' for the dump written on failure (n None,
    all_modules.f90:1096-1097)."""
    tag = "X" * width if n is None else f"{int(n):0{width}d}"
    return f"{simname}{kind}_{tag}.h5"


# ------------------------------------------------------------ primitives
def create(path: str, xyzdims: Sequence[int], time_hours: float = 0.0):
    d = (C.c_int * 3)(*[int(v) for v in xyzdims])
    _check(lib().sparc_h5_create(path.encode(), d, float(time_hours)), f"create {path}")


def attrs(path: str):
    d = (C.c_int * 3)()
    t = C.c_double()
    _check(lib().sparc_h5_attrs(path.encode(), d, C.byref(t)), f"attributes of {path}")
    return tuple(d), t.value


def has(path: str, name: str) -> bool:
    return _check(lib().sparc_h5_has(path.encode(), name.encode()), f"{path}:{name}") == 1


def _fortran_dims(arr: np.ndarray):
    """C-order array of shape (nr, ..., n1) = Fortran a(n1, ..., nr)."""
    return tuple(reversed(arr.shape))


def write(path: str, name: str, arr: np.ndarray):
    a = np.ascontiguousarray(arr, dtype=np.float64)
    dims = (C.c_long * a.ndim)(*_fortran_dims(a))
    _check(lib().sparc_h5_write(path.encode(), name.encode(), a.ndim, dims,
                                a.ctypes.data_as(C.POINTER(C.c_double)), 1.0),
           f"write {path}:{name}")


def read(path: str, name: str, shape: Sequence[int]) -> np.ndarray:
    """Dataset `name` as a C-order array of `shape` (the reversed Fortran
    extents); the file's extents must match."""
    out = np.empty(tuple(shape), dtype=np.float64)
    dims = (C.c_long * out.ndim)(*_fortran_dims(out))
    _check(lib().sparc_h5_read(path.encode(), name.encode(), out.ndim, dims,
                               out.ctypes.data_as(C.POINTER(C.c_double)), 1.0),
           f"read {path}:{name}")
    return out


# ------------------------------------------------------------ bkg.h5
def write_bkg(path: str, cfg, raw: Dict[str, np.ndarray]):
    """A background in the layout READ_IN_HDF5 reads: the raw (physical)
    arrays of ``background.make_background``."""
    create(path, (cfg.nx, cfg.ny, cfg.nz), 0.0)
    names = ["rho0", "pre0", "cs0"] + (["bx0", "by0", "bz0"] if cfg.magnetic == 2 else [])
    for n in names:
        write(path, n, raw[n])
    write(path, "zaxis", raw["zaxis"])
    write(path, "grav", raw["grav"])


def read_bkg(path: str, cfg) -> Dict[str, np.ndarray]:
    """READ_IN_HDF5 (all_modules.f90:179-282): the physical arrays, for
    ``initialise.prepare``; the file's xyzdims must match the run."""
    dims, _ = attrs(path)
    if tuple(dims) != (cfg.nx, cfg.ny, cfg.nz):
        raise H5Error(f"{path}: xyzdims {dims} != run grid {(cfg.nx, cfg.ny, cfg.nz)}")
    shape = (cfg.nz, cfg.ny, cfg.nx)
    out = {n: read(path, n, shape) for n in ("rho0", "pre0", "cs0")}
    if cfg.magnetic == 2:
        for n in ("bx0", "by0", "bz0"):
            out[n] = read(path, n, shape)
    out["zaxis"] = read(path, "zaxis", (cfg.nz,))
    out["grav"] = read(path, "grav", (cfg.nz,))
    return out


# ------------------------------------------------------------ full_NNNN.h5
def _scales(st, nv):
    """Physical unit of each state field: rho, v, v, v, p, b, b, b."""
    return [("rho", 0, lambda x: x * st.dimrho, lambda x: x / st.dimrho),
            ("vx", 1, lambda x: x * st.dimc, lambda x: x / st.dimc),
            ("vy", 2, lambda x: x * st.dimc, lambda x: x / st.dimc),
            ("vz", 3, lambda x: x * st.dimc, lambda x: x / st.dimc),
            ("pre", 4, lambda x: x * st.dimrho * st.dimc ** 2, lambda x: x / st.dimrho / st.dimc ** 2),
            ("bx", 5, lambda x: x * st.dimb, lambda x: x / st.dimb),
            ("by", 6, lambda x: x * st.dimb, lambda x: x / st.dimb),
            ("bz", 7, lambda x: x * st.dimb, lambda x: x / st.dimb)][:nv]


def write_full(path: str, solver, time: int):
    """WRITE_OUT_FULL_STATE: the state in physical units and the CPML bases
    (the RK base values zpmlvars & co., unscaled), time = time index."""
    cfg, st = solver.cfg, solver.st
    a, psi = solver.get_state(with_psi=True)
    create(path, (cfg.nx, cfg.ny, cfg.nz), time * cfg.timestep / 3600.0)
    for name, f, to_file, _ in _scales(st, cfg.nv):
        write(path, name, to_file(a[f]))
    for f, v in sorted(psi.items()):
        write(path, PML_NAMES[f], v)


def read_full(path: str, solver) -> float:
    """READ_IN_FULL_STATE: load a snapshot into the solver (state and CPML
    bases); returns the file's time in hours."""
    cfg, st = solver.cfg, solver.st
    dims, hours = attrs(path)
    if tuple(dims) != (cfg.nx, cfg.ny, cfg.nz):
        raise H5Error(f"{path}: xyzdims {dims} != run grid {(cfg.nx, cfg.ny, cfg.nz)}")
    shape = (cfg.nz, cfg.ny, cfg.nx)
    a = np.empty((cfg.nv,) + shape)
    for name, f, _, from_file in _scales(st, cfg.nv):
        a[f] = from_file(read(path, name, shape))
    psi = {}
    for f in range(6):
        if cfg.face_active(f):
            psi[f] = read(path, PML_NAMES[f], (cfg.npsi,) + tuple(cfg.face_shape(f)))
    solver.put_state(a, psi)
    return hours


# ------------------------------------------------------------ chunk_NNNN.h5
def write_chunk(path: str, solver, time: int, xr=None, yr=None, zr=None):
    """WRITE_OUT_CHUNK over the 0-based inclusive index ranges xr, yr, zr
    (default: everything): axes xaxis*xlength, yaxis*ylength (cm) and
    (zaxis-1)*rsun/1e8 (Mm), the fields in physical units."""
    cfg, st = solver.cfg, solver.st
    xr = xr or (0, cfg.nx - 1)
    yr = yr or (0, cfg.ny - 1)
    zr = zr or (0, cfg.nz - 1)
    sx, sy, sz = (slice(r[0], r[1] + 1) for r in (xr, yr, zr))
    nxc, nyc, nzc = (r[1] - r[0] + 1 for r in (xr, yr, zr))
    create(path, (nxc, nyc, nzc), time * cfg.timestep / 3600.0)
    write(path, "xaxis", st.xaxis[sx] * cfg.xlength)
    write(path, "yaxis", st.yaxis[sy] * cfg.ylength)
    write(path, "zaxis", (st.zaxis[sz] - 1.0) * RSUN / 1.0e8)
    slab = {f: solver.get_slab(f, zr[0], zr[1])[:, sy, sx] for f in range(cfg.nv)}
    order = ("vx", "vy", "vz", "rho", "pre", "bx", "by", "bz")
    sc = {name: (f, to_file) for name, f, to_file, _ in _scales(st, cfg.nv)}
    for name in order:
        if name in sc:
            f, to_file = sc[name]
            write(path, name, to_file(slab[f]))
