"""PULSE = 0 driver on the host: the FITS forcing cube, LAGRANGE_INTERP,
source_dep and source_damp (driver.f90:38-39, 73-80, 130-133;
all_modules.f90:289-301, 470-517, 1892-1950; initialize.f90:1202-1216).

Pins: the FITS layout read back as raw bytes per the FITS standard (2880-byte
header blocks, big-endian IEEE doubles, first axis fastest); 7-point Lagrange
interpolation reproduces any polynomial of degree <= 6 between the cube's
planes and the plane values at its nodes; source_dep peaks at the plane
nearest the requested depth with the Gaussian's exp(-1/4) point at one sigma;
source_damp is 1 for periodic runs and 1 - ((edge - x)/L)^n inside the x/y
CPML (0 at the outer edge, 1 outside the CPML).
"""
import os

import numpy as np
import pytest

from sparc2_amd import forcing as fz
from sparc2_amd.background import make_background
from sparc2_amd.initialise import prepare
from tests.cases import mhd_cfg, quiet_cfg

needs_fits = pytest.mark.skipif(not os.path.exists(fz.LIB), reason="libsparc_fits.so not built")


def _state(cfg):
    """Set-up with the host operators only (no device): prepare + the parts of
    initialise_rhs the forcing needs."""
    from sparc2_amd.initialise import initialise_rhs

    class Ops:  # numpy stand-ins for the set-up derivatives
        def ddx(self, a, bc):
            return np.gradient(a, axis=2)

        def ddy(self, a, bc):
            return np.gradient(a, axis=1)

        def ddz(self, a, bc):
            return np.gradient(a, axis=0)

    bg = make_background(cfg, -5.0, 1.0, R0=6e8) if cfg.mhd else make_background(cfg, -10.0, 0.5)
    st = prepare(cfg, bg)
    initialise_rhs(st, Ops())
    return st


@needs_fits
def test_fits_cube_layout(tmp_path):
    cube = np.arange(5 * 3 * 4, dtype=np.float64).reshape(5, 3, 4) * 0.5 - 3.0   # Fortran (4, 3, 5)
    p = str(tmp_path / "f.fits")
    fz.write_fits_cube(p, cube)
    raw = open(p, "rb").read()
    cards = [raw[i:i + 80].decode() for i in range(0, 2880, 80)]
    kv = {c[:8].strip(): c[10:30].strip() for c in cards if "=" in c[8:10]}
    assert (kv["BITPIX"], kv["NAXIS"], kv["NAXIS1"], kv["NAXIS2"], kv["NAXIS3"]) == ("-64", "3", "4", "3", "5")
    data = np.frombuffer(raw[2880:2880 + cube.size * 8], dtype=">f8").reshape(5, 3, 4)
    np.testing.assert_array_equal(data, cube)
    np.testing.assert_array_equal(fz.read_fits_cube(p, 4, 3, 5), cube)
    with pytest.raises(fz.FitsError):
        fz.read_fits_cube(p, 4, 3, 6)


def test_num_steps():
    cfg = quiet_cfg(pulse=False, timestep=1.0, solartime=500.0 / 3600.0, cadforcing=60.0)
    assert cfg.maxtime == 500
    assert fz.num_steps(cfg) == 12          # floor(500/60) + 2 = 10 -> 2*(10//2) + 2
    cfg = cfg.replace(cadforcing=45.0)      # floor(500/45) + 2 = 13 -> 2*6 + 2
    assert fz.num_steps(cfg) == 14


def test_lagrange_reproduces_degree_six_polynomials():
    cfg = quiet_cfg(pulse=False, timestep=1.0, cadforcing=10.0, randomexciteampl=1.0)
    st = _state(cfg)
    cad, nsteps = cfg.cadforcing, 16
    rng = np.random.default_rng(1)
    coef = rng.standard_normal((7, cfg.ny, cfg.nx))
    P = lambda t: sum(coef[d] * (t / 50.0) ** d for d in range(7))
    cube = np.stack([P((p - 1) * cad) for p in range(1, nsteps + 1)])   # plane p at t = (p-1)*cad
    s = fz.ForcingSeries(st, cube)
    scale = cfg.randomexciteampl / st.dimc
    # interior: all seven terms active (time1 >= 4 and time1 < nsteps - 1)
    for time in (30, 31, 37, 44, 59, 81, 119):
        got = s.lagrange(time)
        np.testing.assert_allclose(got, scale * P(float(time)), rtol=1e-9, atol=1e-12 * scale)
    # at a node the interpolant is the plane itself
    s2 = fz.ForcingSeries(st, cube)
    np.testing.assert_allclose(s2.lagrange(60), scale * cube[6], rtol=1e-12)
    # the cached weights give what a fresh series gives
    s3, s4 = fz.ForcingSeries(st, cube), fz.ForcingSeries(st, cube)
    for t in range(30, 50):
        a = s3.lagrange(t)
    np.testing.assert_array_equal(a, s4.lagrange(49))


def test_source_dep_and_damp():
    cfg = mhd_cfg(pulse=False, randomexcitedepth=-0.3e8, randomexcitesigma=0.2e8,
                  npml=(6, 6, 5, 5, 7, 7))
    st = _state(cfg)
    from sparc2_amd.config import RSUN
    z = st.zaxis
    sd = fz.source_dep(st)
    e = int(np.argmin(np.abs(z - (1.0 - 0.3e8 / RSUN))))
    assert sd[e] == 1.0 and sd.argmax() == e
    dz = (z - z[e]) / (0.2e8 / RSUN)
    np.testing.assert_allclose(sd, np.exp(-0.25 * dz ** 2), rtol=1e-12)
    d = st.source_damp
    assert d.shape == (cfg.ny, cfg.nx)
    assert d[0, cfg.nx // 2] == 0.0 and d[cfg.ny // 2, 0] == 0.0          # outer edges
    assert d[cfg.ny // 2, cfg.nx // 2] == 1.0                            # interior
    assert d[cfg.ny // 2, 5] == 1.0 and d[cfg.ny // 2, cfg.nx - 6] == 1.0  # inner CPML edge, s = 0
    assert np.all((d >= 0) & (d <= 1))
    st2 = _state(quiet_cfg(pulse=False))
    np.testing.assert_array_equal(st2.source_damp, 1.0)
