"""Shared run configurations and seeded states for the tests."""
from __future__ import annotations

import numpy as np

from sparc2_amd.background import make_background
from sparc2_amd.config import RunConfig


def quiet_cfg(nx=32, ny=32, nz=40, **kw):
    base = dict(nx=nx, ny=ny, nz=nz, xlength=25e8, ylength=25e8, timestep=1.0,
                npml=(0, 0, 0, 0, 8, 8), pmlc=((9.2103, 2.0, 1.0, 0.1),) * 6,
                pulse_px=12.5e8, pulse_py=12.5e8, pulse_t1=20.0, pulse_st=8.0,
                pulse_t0=40.0, pulse_amp=1.0e4, t_z=3.0, t_xy=5.0)
    base.update(kw)
    return RunConfig(**base)


def mhd_cfg(nx=32, ny=32, nz=40, periodic=False, **kw):
    base = dict(nx=nx, ny=ny, nz=nz, xlength=30e8, ylength=30e8, timestep=0.25, magnetic=2,
                periodic=periodic, use_hpml=not periodic,
                npml=(6, 6, 6, 6, 7, 7) if not periodic else (0, 0, 0, 0, 7, 7),
                pmlc=((9.2103, 2.0, 1.0, 0.1),) * 6, pulse_px=19e8, pulse_py=15e8,
                pulse_pz=-0.5e8, pulse_sr=2e8, pulse_sz=0.5e8,
                pulse_t1=5.0, pulse_st=2.0, pulse_t0=10.0, t_z=1.0, t_xy=1.5, maxva=1.0e7)
    base.update(kw)
    return RunConfig(**base)


def background(cfg, zbot=None, ztop=None):
    if cfg.magnetic == 2:
        return make_background(cfg, -5.0 if zbot is None else zbot, 1.0 if ztop is None else ztop,
                               R0=6e8)
    bg = make_background(cfg, -10.0 if zbot is None else zbot, 0.5 if ztop is None else ztop)
    if cfg.magnetic == 1:
        # thermal: give the column a horizontal structure so 3-D backgrounds matter
        nz, ny, nx = bg["rho0"].shape
        x = np.arange(nx) / nx
        y = np.arange(ny) / ny
        mod = 1.0 + 0.05 * np.sin(2 * np.pi * x)[None, None, :] * np.cos(2 * np.pi * y)[None, :, None]
        bg["rho0"] = np.ascontiguousarray(bg["rho0"] * mod)
        bg["pre0"] = np.ascontiguousarray(bg["pre0"] * mod)
    return bg


def random_state(cfg, seed=1, amp=1e-4):
    """Seeded smooth-ish perturbation state and PML bases.  Each field is
    scaled by its local background (rho by rho0, p by p0, v by the sound
    speed, b by sqrt(rho0) c, all in the units of the background at cell
    (1,1,1)), so rho+rho0 and p+p0 stay positive in every layer and the
    reference's failure checks (driver.f90:175-183) pass."""
    rng = np.random.default_rng(seed)
    shape = (cfg.nv, cfg.nz, cfg.ny, cfg.nx)
    a = rng.standard_normal(shape)
    # two passes of a 3-point average in each direction to tame grid noise
    for ax in (1, 2, 3):
        for _ in range(2):
            a = 0.25 * np.roll(a, 1, ax) + 0.5 * a + 0.25 * np.roll(a, -1, ax)
    bg = background(cfg)
    r0, c0 = bg["rho0"][0, 0, 0], bg["cs0"][0, 0, 0]
    rho = bg["rho0"] / r0
    c = bg["cs0"] / c0
    p = bg["pre0"] / (r0 * c0 * c0)
    scale = [rho, c, c, c, p] + [np.sqrt(rho) * c] * 3
    for q in range(cfg.nv):
        a[q] *= amp * scale[q]
    psi = {}
    for f in range(6):
        if cfg.face_active(f):
            psi[f] = amp * 0.1 * rng.standard_normal((cfg.npsi,) + cfg.face_shape(f))
    return np.ascontiguousarray(a), psi


def rel_l2(x, ref):
    n = np.linalg.norm(ref.ravel())
    d = np.linalg.norm((x - ref).ravel())
    return d / n if n > 0 else d


def field_errors(x, ref):
    return [rel_l2(x[f], ref[f]) for f in range(x.shape[0])]
