"""Synthetic solar backgrounds in the reference's input format (cgs).

The reference reads a background file holding 3-D ``rho0``, ``pre0``, ``cs0``
(and ``bx0``, ``by0``, ``bz0`` for MHD) plus 1-D ``zaxis`` (Mm) and ``grav``
(cgs) (all_modules.f90:179-275).  No such file ships with the reference, so the
workloads are generated here, deterministically:

* ``polytrope_column`` -- adiabatic polytrope of index m = 1.5 (Gamma = 5/3)
  scaled to rho = 2.8e-7 g/cm^3 and c = 7 km/s at z = 0.  Its surface lies at
  z = c0^2 (m+1) / (Gamma g) = +0.268 Mm, inside a domain that reaches +0.5 Mm,
  so above z = 0 the column continues as an isothermal layer (c = c0, scale
  height H = c0^2 / (Gamma g) = 107 km), continuous in rho, p, c and dp/dz.
* ``flux_tube`` -- a self-similar axisymmetric tube, Gaussian in radius,
  B_z = B0 (R0/R)^2 exp(-r^2/R^2) with R(z)^2 = R0^2 (1 + exp(z/H_R)) / 2, and
  B_r = r R'/R B_z (divergence free for any R(z)).  Gas pressure and density are
  reduced by horizontal magnetohydrostatic balance, p = p_ext - B^2/(8 pi),
  floored at ``p_floor`` * p_ext where the field would exceed the external
  pressure, with the temperature (sound speed) of the external column.  The
  floor (default 0.5) bounds the density contrast across the tube wall in
  the low-beta layer above the photosphere: there the tube is not in
  magnetohydrostatic balance (a 3 kG field cannot be confined by the gas
  pressure), and with a floor of 0.05 the linearised equations around it
  grow within a few hundred steps.

Arrays are returned with numpy shape (nz, ny, nx), C order, which is the
reference's Fortran (nx, ny, nz) memory layout.
"""
from __future__ import annotations

import numpy as np

from .config import RunConfig

G_SUN = 2.74e4


def polytrope_column(zaxis_mm, rho_s=2.8e-7, c_s=7.0e5, g=G_SUN, m=1.5, gamma=5.0 / 3.0):
    z = np.asarray(zaxis_mm, dtype=np.float64) * 1.0e8
    p_s = c_s * c_s * rho_s / gamma
    d0 = c_s * c_s * (m + 1.0) / (gamma * g)      # depth of z = 0 below the surface
    H = c_s * c_s / (gamma * g)
    below = z <= 0.0
    ratio = np.where(below, (d0 - z) / d0, 1.0)
    rho = np.where(below, rho_s * ratio ** m, rho_s * np.exp(-z / H))
    p = np.where(below, p_s * ratio ** (m + 1.0), p_s * np.exp(-z / H))
    c = np.where(below, c_s * np.sqrt(ratio), c_s)
    return rho, p, c


def horizontal_axes(cfg: RunConfig):
    """Physical x, y positions (cm) of the grid points (initialize.f90:470-492)."""
    if cfg.periodic:
        x = np.arange(cfg.nx) / cfg.nx * cfg.xlength
        y = np.arange(cfg.ny) / cfg.ny * cfg.ylength
    else:
        x = np.arange(cfg.nx) / (cfg.nx - 1) * cfg.xlength
        y = np.arange(cfg.ny) / (cfg.ny - 1) * cfg.ylength
    return x, y


def quiet_background(cfg: RunConfig, zaxis_mm, g=G_SUN):
    rho, p, c = polytrope_column(zaxis_mm, g=g)
    shape = (cfg.nz, cfg.ny, cfg.nx)
    bc = lambda v: np.ascontiguousarray(np.broadcast_to(v[:, None, None], shape))
    return {
        "rho0": bc(rho), "pre0": bc(p), "cs0": bc(c),
        "zaxis": np.asarray(zaxis_mm, dtype=np.float64).copy(),
        "grav": np.full(cfg.nz, g, dtype=np.float64),
    }


def flux_tube(cfg: RunConfig, zaxis_mm, B0=3000.0, R0=4.0e8, H_R=1.0e8, center=None,
              p_floor=0.5, g=G_SUN):
    rho_e, p_e, c_e = polytrope_column(zaxis_mm, g=g)
    x, y = horizontal_axes(cfg)
    xc, yc = center if center is not None else (0.5 * cfg.xlength, 0.5 * cfg.ylength)
    z = np.asarray(zaxis_mm, dtype=np.float64) * 1.0e8
    ez = np.exp(z / H_R)
    R2 = R0 * R0 * (1.0 + ez) / 2.0                  # (nz,)
    dlnR = 0.5 * (ez / H_R) / (1.0 + ez)             # R'/R
    dx = (x - xc)[None, None, :]
    dy = (y - yc)[None, :, None]
    r2 = dx * dx + dy * dy
    bz = B0 * (R0 * R0 / R2)[:, None, None] * np.exp(-r2 / R2[:, None, None])
    bx = dx * dlnR[:, None, None] * bz
    by = dy * dlnR[:, None, None] * bz
    pm = (bx * bx + by * by + bz * bz) / (8.0 * np.pi)
    pe = p_e[:, None, None]
    p = np.maximum(pe - pm, p_floor * pe)
    rho = rho_e[:, None, None] * (p / pe)
    c = np.broadcast_to(c_e[:, None, None], p.shape)
    return {
        "rho0": np.ascontiguousarray(rho), "pre0": np.ascontiguousarray(p),
        "cs0": np.ascontiguousarray(c),
        "bx0": np.ascontiguousarray(bx), "by0": np.ascontiguousarray(by),
        "bz0": np.ascontiguousarray(bz),
        "zaxis": np.asarray(zaxis_mm, dtype=np.float64).copy(),
        "grav": np.full(cfg.nz, g, dtype=np.float64),
    }


def make_background(cfg: RunConfig, zbot_mm: float, ztop_mm: float, **kw):
    """Background for ``cfg``: quiet polytrope (magnetic 0/1) or flux tube (2)."""
    zaxis = np.linspace(zbot_mm, ztop_mm, cfg.nz)
    if cfg.magnetic == 2:
        return flux_tube(cfg, zaxis, **kw)
    return quiet_background(cfg, zaxis)
