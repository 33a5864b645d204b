"""Host-side FP64 set-up of a run: everything the time step reads that the
reference computes once before the time loop.

* ``nondimensionalise``  -- READ_IN_HDF5 after the read (all_modules.f90:284-412)
* ``grid_metric``        -- axes, stretchx/y, unstretch/stretch, deltat
                            (initialize.f90:468-517, 674-733)
* ``rk_betas``           -- INITIALIZE_STEP (initialize.f90:1366-1381)
* ``initialise_rhs``     -- INITIALISE_RHS (initialize.f90:883-1228): sponges,
                            Alfven limiter, curl B0, rhoinv, c2rho0, background
                            gradients, STABILISE, CPML profiles
* ``source_profiles`` / ``source_amplitude`` -- the pulse of driver.f90:137-145
                            in separable form

The derivatives the set-up needs (curl B0, background gradients) are taken
from an operator object with ``ddx/ddy/ddz(array, bc)``; the solver passes the
GPU drop-in operators.  Arrays are numpy (nz, ny, nx), C order.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, Optional

import numpy as np

from .config import PI, RSUN, RunConfig

F32_0_9998 = float(np.float32(0.9998))   # default-real literals of initialize.f90:1023-1028
F32_1_0001 = float(np.float32(1.0001))

# centred and one-sided DRP rows (initialize.f90:1959-2029); row 5 = centred
DPZ = np.array([
    [-2.391602219538, 5.832490322294, -7.650218001182 + 0.000000000001, 7.907810563576,
     -5.922599052629, 3.071037015445, -1.014956769726, 0.170022256519, 0.002819958377,
     -0.004791009708, -0.000013063429],
    [-0.180022054228, -1.237550583044, 2.484731692990, -1.810320814061, 1.112990048440,
     -0.481086916514, 0.126598690230, -0.015510730165, 0.000024609059 - 0.000003,
     0.000156447571 - 0.000000000001, -0.000007390277],
    [0.057982271137, -0.536135360383, -0.264089548967 + 0.000000000002,
     0.917445877606 - 0.000000000002, -0.169688364841, -0.029716326170, 0.029681617641,
     -0.005222483773, -0.000118806260, -0.000118806260, -0.000020069730],
    [-0.013277273810, 0.115976072920, -0.617479187931, -0.274113948206 + 0.000000000002,
     1.086208764655 - 0.000000000002, -0.402951626982, 0.131066986242, -0.028154858354,
     0.002596328316, 0.000128743150, 0.0],
    [0.016756572303, -0.117478455239, 0.411034935097, -1.130286765151,
     0.341435872100 - 0.000000000001, 0.556396830543, -0.082525734207, 0.003565834658,
     0.001173034777, -0.000071772671 + 0.000000000064, -0.000000352273],
    [-0.002484594688, +0.020779405824, -0.090320001280, +0.286511173973, -0.872756993962, 0.0,
     +0.872756993962, -0.286511173973, +0.090320001280, -0.020779405824, +0.002484594688],
])


@dataclasses.dataclass
class SetupState:
    cfg: RunConfig
    dimc: float
    dimrho: float
    diml: float
    dimb: float
    xaxis: np.ndarray
    yaxis: np.ndarray
    zaxis: np.ndarray          # nondimensional, 1 + z/R
    stretchx: float
    stretchy: float
    stretch: np.ndarray
    deltat: float
    betas: np.ndarray
    bg: Dict[str, np.ndarray]  # nondimensional 3-D/1-D arrays
    pml: Dict[int, Dict[str, np.ndarray]] = dataclasses.field(default_factory=dict)
    # PULSE = 0 driver: source_damp(x, y) of initialize.f90:1202-1216, C (ny, nx)
    source_damp: Optional[np.ndarray] = None
    bg3d: bool = True


def nondimensionalise(cfg: RunConfig, raw: Dict[str, np.ndarray]):
    """all_modules.f90:284-412 applied to the arrays of a background file."""
    shape = (cfg.nz, cfg.ny, cfg.nx)
    rho0 = np.array(raw["rho0"], dtype=np.float64).reshape(shape)
    p0 = np.array(raw["pre0"], dtype=np.float64).reshape(shape)
    cs = np.array(raw["cs0"], dtype=np.float64).reshape(shape)
    g = np.array(raw["grav"], dtype=np.float64).copy()
    zaxis = 1.0 + np.asarray(raw["zaxis"], dtype=np.float64) * 1.0e8 / RSUN
    dimc = float(cs[0, 0, 0])
    dimrho = float(rho0[0, 0, 0])
    diml = RSUN
    dimb = (4.0 * PI * (dimc * dimc) * dimrho) ** 0.5
    cq = cs[:, 0, 0].copy()
    if cfg.magnetic == 0:          # QUIET: horizontal copy of column (1,1,:)
        rho0 = np.ascontiguousarray(np.broadcast_to(rho0[:, :1, :1], shape))
        p0 = np.ascontiguousarray(np.broadcast_to(p0[:, :1, :1], shape))
        cs = np.ascontiguousarray(np.broadcast_to(cs[:, :1, :1], shape))
    g = g * diml / (dimc * dimc)
    cq = cq / dimc
    p0 = p0 / (dimrho * (dimc * dimc))
    rho0 = rho0 / dimrho
    cs = cs / dimc
    out = {"rho0": rho0, "p0": p0, "c_speed": cs, "c2": cs * cs, "g": g, "cq": cq}
    if cfg.magnetic == 2:
        for src, dst in (("bx0", "box"), ("by0", "boy"), ("bz0", "boz")):
            out[dst] = np.array(raw[src], dtype=np.float64).reshape(shape) / dimb
    return out, zaxis, dimc, dimrho, diml, dimb


def grid_metric(cfg: RunConfig, zaxis: np.ndarray):
    """Axes and metric factors, initialize.f90:468-517 and 674-730."""
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    if cfg.periodic:
        xaxis = np.arange(nx, dtype=np.float64) / float(nx)
        yaxis = np.arange(ny, dtype=np.float64) / float(ny)
        stretchx = RSUN / cfg.xlength * nx
        stretchy = RSUN / cfg.ylength * ny
    else:
        xaxis = np.arange(nx, dtype=np.float64) / float(nx - 1)
        yaxis = np.arange(ny, dtype=np.float64) / float(ny - 1)
        stretchx = RSUN / cfg.xlength * (nx - 1)
        stretchy = RSUN / cfg.ylength * (ny - 1)
    z = zaxis
    uns = np.zeros(nz)
    for k in range(nz):
        if k < 5:
            acc = 0.0
            for m in range(11):
                acc += z[m] * DPZ[k, m]
            uns[k] = acc * (nz - 1)
        elif k >= nz - 5:
            r = nz - 1 - k
            acc = 0.0
            for m in range(11):
                acc += z[nz - 1 - m] * DPZ[r, m]
            uns[k] = -acc * (nz - 1)
        else:
            uns[k] = ((-z[k - 5] + z[k + 5]) * DPZ[5, 10] + (-z[k - 4] + z[k + 4]) * DPZ[5, 9]
                      + (-z[k - 3] + z[k + 3]) * DPZ[5, 8] + (-z[k - 2] + z[k + 2]) * DPZ[5, 7]
                      + (-z[k - 1] + z[k + 1]) * DPZ[5, 6]) * (nz - 1)
    stretch = (nz - 1) * 1.0 / uns
    return xaxis, yaxis, stretchx, stretchy, stretch


def rk_betas(deltat: float) -> np.ndarray:
    opt = [1.0, 0.5, 0.1665579725791151184, 0.03950410250586729987, 0.007810706393337838236]
    b = [0.0] * 5
    b[4] = opt[1]
    b[3] = opt[2] / b[4]
    b[2] = opt[3] / (b[4] * b[3])
    b[1] = opt[4] / (b[4] * b[3] * b[2])
    return np.array(b) * deltat


def _reduction_search(bg, dimb, dimrho, maxva, ops=None):
    """Decimal search of the Alfven limiter, initialize.f90:962-986.  With an
    operator object that offers ``alfven_search`` (the GPU solver) the
    maximum over the grid is evaluated on the device."""
    box, boy, boz, rho0 = bg["box"], bg["boy"], bg["boz"], bg["rho0"]
    cq = bg["cq"][:, None, None]
    red = (boz * boz + box * box + boy * boy) / (rho0 * (cq * cq))
    b2 = box * box + boy * boy + boz * boz
    if ops is not None and hasattr(ops, "alfven_search"):
        temp = ops.alfven_search(red, b2, rho0, dimb, dimrho, maxva)
        return temp / (temp + red)
    den = 4.0 * PI * dimrho * rho0
    va, dtemp = 0.0, 1.0e9
    temp = -dtemp
    tmp = np.empty_like(red)
    if np.nanmax(red) > 0:
        while dtemp > 1.0e-5 and temp < 1.0e10 - 1.1:
            while va < maxva and temp < 1.0e10:
                temp = temp + dtemp
                with np.errstate(invalid="ignore", divide="ignore"):
                    np.add(temp, red, out=tmp)
                    np.divide(dimb * dimb * temp, tmp, out=tmp)
                    np.multiply(tmp, b2, out=tmp)
                    np.divide(tmp, den, out=tmp)
                    np.sqrt(tmp, out=tmp)
                va = float(np.nanmax(tmp))
            temp = temp - dtemp
            dtemp = dtemp / 10.0
            va = 0.0
    else:
        temp = 1.0e10
    return temp / (temp + red)


def initialise_rhs(st: SetupState, ops) -> None:
    """INITIALISE_RHS (initialize.f90:883-1228) on the nondimensional state."""
    cfg, bg = st.cfg, st.bg
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    shape = (nz, ny, nx)
    z = st.zaxis
    bc = 5 if cfg.periodic else 1
    # sponges (initialize.f90:899-947); spongez is left undefined by the
    # reference when USE_PML is on -- its intended value 0 is used
    if cfg.use_pml:
        spongez = np.zeros(nz)
    else:
        spongez = (400.0 / (1.0 + np.exp((1.002 - z) / 0.0001))
                   + 40.0 / (1.0 + np.exp((z - (z[0] + 0.005)) / 0.0005)))
    if cfg.periodic:
        sponge = np.ascontiguousarray(np.broadcast_to(spongez[:, None, None], shape))
    else:
        spx = np.zeros(nx)
        spy = np.zeros(ny)
        if not cfg.use_hpml:
            s0 = cfg.sp_length / cfg.xlength
            s1 = 1.0 - cfg.sp_length / cfg.xlength
            spx = cfg.sp_str * (1.0 - 1.0 / ((1.0 + np.exp((s0 - st.xaxis) / cfg.sp_decay))
                                             * (1.0 + np.exp((st.xaxis - s1) / cfg.sp_decay))))
            spy = cfg.sp_str * (1.0 - 1.0 / ((1.0 + np.exp((s0 - st.yaxis) / cfg.sp_decay))
                                             * (1.0 + np.exp((st.yaxis - s1) / cfg.sp_decay))))
        sponge = spx[None, None, :] + spy[None, :, None] + spongez[:, None, None]
        sponge = np.ascontiguousarray(sponge)

    if cfg.mhd:
        bg["reduction"] = _reduction_search(bg, st.dimb, st.dimrho, cfg.maxva, ops)
        t = ops.ddy(bg["box"], bc)
        bg["curlboz"] = ops.ddx(bg["boy"], bc) - t
        t = ops.ddy(bg["boz"], bc)
        bg["curlbox"] = t - ops.ddz(bg["boy"], 1)
        t = ops.ddx(bg["boz"], bc)
        bg["curlboy"] = ops.ddz(bg["box"], 1) - t

    rho0 = bg["rho0"]
    c2 = bg["c2"]
    g = bg["g"].copy()
    bg["rhoinv"] = 1.0 / rho0
    bg["c2rho0"] = c2 * rho0
    grz = ops.ddz(rho0, 1)
    gpz = ops.ddz(bg["p0"], 1)
    if cfg.stabilise:   # initialize.f90:1018-1034
        zz = z[:, None, None]
        v2 = -rho0 * g[:, None, None]
        near = (zz > F32_0_9998) & (zz < 1.0)
        rest = (zz <= F32_0_9998) | (zz >= F32_1_0001)
        v1n = c2 * grz + 1.e-4 * RSUN / (st.dimrho * (st.dimc * st.dimc))
        v1r = c2 * grz
        gpz = np.where(near, np.maximum(v1n, v2), np.where(rest, np.maximum(v1r, v2), gpz))

    if cfg.use_pml:
        if cfg.mhd:
            b2 = bg["box"] ** 2 + bg["boy"] ** 2 + bg["boz"] ** 2
            tin = np.sqrt(c2 + bg["reduction"] * b2 / rho0)
        else:
            tin = np.sqrt(c2)
        f0 = PI * 0.005 * st.diml / st.dimc
        # z faces (initialize.f90:1048-1104), alpha-scaling of the gradients and g
        for f in (4, 5):
            if not cfg.face_active(f):
                continue
            n = cfg.npml[f]
            RC, pn, kap_max, sp = cfg.pmlc[f]
            if f == 4:
                L = z[n - 1] - z[0]
                ks = np.arange(n)
                s = (z[n - 1] - z[ks]) / L
                alpha = f0 * (z[ks] - z[0]) / L
                tsl = tin[n - 1]
            else:
                k0 = nz - n
                L = z[nz - 1] - z[k0]
                ks = k0 + np.arange(n)
                s = -(z[k0] - z[ks]) / L
                alpha = f0 * (z[nz - 1] - z[ks]) / L
                tsl = tin[k0]
            sn = s ** pn
            damp = RC * tsl[None, :, :] * (pn + 1.0) / (2.0 * L) * sn[:, None, None]
            kap = np.broadcast_to((1.0 + (kap_max - 1.0) * sn)[:, None, None], damp.shape)
            al = alpha[:, None, None]
            st.pml[f] = {"b": -al - damp / kap, "a": -damp / (kap * kap),
                         "kappa": np.ascontiguousarray(kap)}
            fac = al / kap / (damp / kap + al)
            grz[ks] = grz[ks] * al / kap / (damp / kap + al)
            gpz[ks] = gpz[ks] * al / kap / (damp / kap + al)
            d11, k11 = damp[:, 0, 0], kap[:, 0, 0]
            g[ks] = g[ks] * alpha / k11 / (d11 / k11 + alpha)
            sponge[ks] = sponge[ks] + sp * tsl[None, :, :] * (s ** (pn + 1.0))[:, None, None]
            del fac
        # source_dampx / source_dampy: the x/y CPML profiles ((edge - x)/L)^n
        sdx, sdy = np.zeros(nx), np.zeros(ny)
        if cfg.use_hpml:
            xa, ya = st.xaxis, st.yaxis
            for f in (0, 1):   # initialize.f90:1108-1152
                if not cfg.face_active(f):
                    continue
                n = cfg.npml[f]
                RC, pn, kap_max, sp = cfg.pmlc[f]
                if f == 0:
                    L = xa[n - 1] - xa[0]
                    iis = np.arange(n)
                    s = (xa[n - 1] - xa[iis]) / L
                    alpha = f0 * (xa[iis] - xa[0]) / L
                    tsl = tin[:, :, n - 1]
                else:
                    i0 = nx - n
                    L = xa[nx - 1] - xa[i0]
                    iis = i0 + np.arange(n)
                    s = -(xa[i0] - xa[iis]) / L
                    alpha = f0 * (xa[nx - 1] - xa[iis]) / L
                    tsl = tin[:, :, i0]
                sn = s ** pn
                sdx[iis] = sn
                damp = RC * tsl[:, :, None] * (pn + 1.0) / (2.0 * L) * sn[None, None, :]
                kap = np.broadcast_to((1.0 + (kap_max - 1.0) * sn)[None, None, :], damp.shape)
                al = alpha[None, None, :]
                st.pml[f] = {"b": -al - damp / kap, "a": -damp / (kap * kap),
                             "kappa": np.ascontiguousarray(kap)}
                sponge[:, :, iis] = sponge[:, :, iis] + sp * tsl[:, :, None] * (s ** (pn + 1.0))[None, None, :]
            for f in (2, 3):   # initialize.f90:1154-1199
                if not cfg.face_active(f):
                    continue
                n = cfg.npml[f]
                RC, pn, kap_max, sp = cfg.pmlc[f]
                if f == 2:
                    L = ya[n - 1] - ya[0]
                    jjs = np.arange(n)
                    s = (ya[n - 1] - ya[jjs]) / L
                    alpha = f0 * (ya[jjs] - ya[0]) / L
                    tsl = tin[:, n - 1, :]
                else:
                    j0 = ny - n
                    L = ya[ny - 1] - ya[j0]
                    jjs = j0 + np.arange(n)
                    s = (-(ya[j0] - ya[jjs])) / L
                    alpha = f0 * (ya[ny - 1] - ya[jjs]) / L
                    tsl = tin[:, j0, :]
                sn = s ** pn
                sdy[jjs] = sn
                damp = RC * tsl[:, None, :] * (pn + 1.0) / (2.0 * L) * sn[None, :, None]
                kap = np.broadcast_to((1.0 + (kap_max - 1.0) * sn)[None, :, None], damp.shape)
                al = alpha[None, :, None]
                st.pml[f] = {"b": -al - damp / kap, "a": -damp / (kap * kap),
                             "kappa": np.ascontiguousarray(kap)}
                sponge[:, jjs, :] = sponge[:, jjs, :] + sp * tsl[:, None, :] * (s ** (pn + 1.0))[None, :, None]

    # source_damp (initialize.f90:1202-1216): 1 when periodic, else one minus
    # the larger x/y CPML profile (1 outside the x/y CPML, 1 without it)
    if cfg.periodic:
        st.source_damp = np.ones((ny, nx))
    elif cfg.use_pml:
        st.source_damp = 1.0 - np.maximum(sdx[None, :], sdy[:, None])
    else:
        st.source_damp = np.ones((ny, nx))
    bg["gradrho0_z"] = grz
    bg["gradp0_z"] = gpz
    bg["g"] = g
    bg["spongexyz"] = sponge
    if cfg.mhd:   # initialize.f90:1219-1228
        bg["gradrho0_x"] = ops.ddx(rho0, bc)
        bg["gradp0_x"] = ops.ddx(bg["p0"], bc)
        bg["gradrho0_y"] = ops.ddy(rho0, bc)
        bg["gradp0_y"] = ops.ddy(bg["p0"], bc)


def prepare(cfg: RunConfig, raw: Dict[str, np.ndarray]) -> SetupState:
    """Everything up to (not including) INITIALISE_RHS."""
    cfg.validate()
    bg, zaxis, dimc, dimrho, diml, dimb = nondimensionalise(cfg, raw)
    xaxis, yaxis, sx, sy, stretch = grid_metric(cfg, zaxis)
    deltat = cfg.timestep * dimc / diml
    bg3d = not (cfg.magnetic == 0 and cfg.periodic and not cfg.use_hpml)
    return SetupState(cfg=cfg, dimc=dimc, dimrho=dimrho, diml=diml, dimb=dimb, xaxis=xaxis,
                      yaxis=yaxis, zaxis=zaxis, stretchx=sx, stretchy=sy, stretch=stretch,
                      deltat=deltat, betas=rk_betas(deltat), bg=bg, bg3d=bg3d)


def damping_rates(st: SetupState) -> np.ndarray:
    """damping_rates of INIT_DAMPING (initialize.f90:1286-1314): per
    horizontal wavenumber (kx = 0..nx/2, ky folded to |ky|), in C order
    (ny, nx/2+1) = the Fortran array (nx/2+1, ny).  The factor 1/(nx*ny)
    normalises the unnormalised R2C/C2R transform pair of damp_velocity."""
    cfg = st.cfg
    nx, ny = cfg.nx, cfg.ny
    constx = 2.0 * PI * RSUN / cfg.xlength
    consty = 2.0 * PI * RSUN / cfg.ylength
    ii = np.arange(nx // 2 + 1, dtype=np.float64)
    jj = np.arange(ny, dtype=np.float64)
    ky = np.where(np.arange(ny) <= ny // 2, jj, ny - jj)
    kay2d = (constx ** 2 * ii[None, :] ** 2 + consty ** 2 * ky[:, None] ** 2) ** 0.5
    kayref = 902.0
    return st.diml / st.dimc / nx / ny * 1.0e-6 * (500.0 * (kay2d / kayref) ** 2.2) * 20.0


def source_profiles(st: SetupState):
    """Separable factors of forcing_p (driver.f90:143-145)."""
    cfg, diml = st.cfg, st.diml
    X = st.xaxis * cfg.xlength / RSUN - cfg.pulse_px / diml
    Y = st.yaxis * cfg.ylength / RSUN - cfg.pulse_py / diml
    Z = st.zaxis - (cfg.pulse_pz / diml + 1.0)
    srd = cfg.pulse_sr / diml
    szd = cfg.pulse_sz / diml
    sx = np.exp(-1.0 * (X * X) / (2.0 * (srd * srd)))
    sy = np.exp(-1.0 * (Y * Y) / (2.0 * (srd * srd)))
    sz = np.exp(-1.0 * (Z * Z) / (2.0 * (szd * szd)))
    return sx, sy, sz


def source_amplitude(st: SetupState, time: int) -> float:
    """Time factor of forcing_p at t = time*timestep (driver.f90:140-142)."""
    cfg = st.cfg
    if not cfg.pulse:
        return 0.0
    t = time * cfg.timestep
    s = math.sin(2.0 * PI * (t - cfg.pulse_t1 + 2.0 * cfg.pulse_t0) / cfg.pulse_t0)
    e = math.exp(-1.0 * (t - cfg.pulse_t1) ** 2 / (2.0 * (cfg.pulse_st * cfg.pulse_st)))
    return cfg.pulse_amp / 0.56458272 / st.dimc * s * e


def layout_for_device(st: SetupState) -> Dict[str, np.ndarray]:
    """Background arrays as the C ABI expects them (z profiles when !bg3d)."""
    out = {}
    names = ["rho0", "rhoinv", "c2rho0", "gradrho0_z", "gradp0_z", "spongexyz", "p0"]
    for n in names:
        v = np.ascontiguousarray(st.bg[n], dtype=np.float64)
        out[n] = np.ascontiguousarray(v[:, 0, 0]) if not st.bg3d else v
    out["g"] = np.ascontiguousarray(st.bg["g"])
    if st.cfg.mhd:
        for n in ("box", "boy", "boz", "curlbox", "curlboy", "curlboz", "gradrho0_x",
                  "gradrho0_y", "gradp0_x", "gradp0_y", "reduction"):
            out[n] = np.ascontiguousarray(st.bg[n], dtype=np.float64)
    return out
