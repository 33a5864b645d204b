"""CPU tests of the host side: configuration, backgrounds, set-up, the ABI."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle.oracle import Oracle
from sparc2_amd import _lib
from sparc2_amd.background import polytrope_column
from sparc2_amd.config import RunConfig
from sparc2_amd.initialise import initialise_rhs, prepare, source_amplitude
from tests.cases import background, mhd_cfg, quiet_cfg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_derived_cadences_and_flags():
    c = RunConfig(nx=32, ny=32, nz=32, xlength=1e9, ylength=1e9, timestep=0.75, t_xy=15.0, t_z=7.5,
                  solartime=4.0)
    assert (c.xyfreq, c.zfreq) == (20, 10)           # floor(t/dt + 1e-5)
    assert c.maxtime == 2 * ((4 * 3600 * 4 // 3) // 2)
    c2 = RunConfig(nx=32, ny=32, nz=32, xlength=1e9, ylength=1e9, timestep=2.0, t_xy=1.0, t_z=0.5)
    assert (c2.xyfreq, c2.zfreq) == (1, 1)            # t < timestep is clamped
    c3 = RunConfig(nx=32, ny=32, nz=32, xlength=1e9, ylength=1e9, use_hpml=True,
                   npml=(4, 4, 4, 4, 6, 6))
    assert c3.periodic is False                       # USE_HPML switches PERIODIC off
    assert RunConfig(nx=32, ny=32, nz=32, xlength=1e9, ylength=1e9).periodic is True
    with pytest.raises(ValueError, match="PERIODIC and USE_HPML"):
        RunConfig(nx=32, ny=32, nz=32, xlength=1e9, ylength=1e9, use_hpml=True, periodic=True)
    c4 = RunConfig(nx=32, ny=32, nz=32, xlength=1e9, ylength=1e9, use_hpml=False,
                   npml=(4, 4, 4, 4, 6, 6))
    assert c4.npml == (0, 0, 0, 0, 6, 6)
    with pytest.raises(ValueError):
        RunConfig(nx=8, ny=32, nz=32, xlength=1e9, ylength=1e9).validate()


def test_polytrope_is_hydrostatic_and_matches_its_anchor():
    z = np.linspace(-10, 0.5, 20001)
    rho, p, c = polytrope_column(z)
    r0, p0, c0 = polytrope_column(np.array([0.0]))
    assert r0[0] == pytest.approx(2.8e-7, rel=1e-14) and c0[0] == pytest.approx(7.0e5, rel=1e-14)
    dpdz = np.gradient(p, z * 1e8)
    np.testing.assert_allclose(dpdz[1:-1], -rho[1:-1] * 2.74e4, rtol=2e-3)
    np.testing.assert_allclose(c ** 2, (5.0 / 3.0) * p / rho, rtol=1e-12)


def test_flux_tube_is_divergence_free():
    cfg = mhd_cfg(nx=48, ny=48, nz=48)
    bg = background(cfg)
    x = np.arange(cfg.nx) / (cfg.nx - 1) * cfg.xlength
    y = np.arange(cfg.ny) / (cfg.ny - 1) * cfg.ylength
    z = bg["zaxis"] * 1e8
    div = (np.gradient(bg["bx0"], x, axis=2) + np.gradient(bg["by0"], y, axis=1)
           + np.gradient(bg["bz0"], z, axis=0))
    scale = np.abs(np.gradient(bg["bz0"], z, axis=0)).max()
    assert np.abs(div[2:-2, 2:-2, 2:-2]).max() < 0.05 * scale
    assert np.all(bg["pre0"] > 0) and np.all(bg["rho0"] > 0)


@pytest.mark.parametrize("case", ["quiet", "thermal", "mhd", "mhd_periodic", "hpml"])
def test_product_setup_matches_oracle_setup(case):
    """sparc2_amd.initialise (host FP64 set-up) with the oracle's derivative
    operators reproduces the oracle's own INITIALISE_RHS."""
    cfg = {"quiet": quiet_cfg(), "thermal": quiet_cfg(magnetic=1), "mhd": mhd_cfg(),
           "mhd_periodic": mhd_cfg(periodic=True),
           "hpml": quiet_cfg(periodic=False, use_hpml=True, npml=(5, 6, 6, 5, 8, 7))}[case]
    bg = background(cfg)
    o = Oracle(cfg, bg)
    st = prepare(cfg, bg)
    initialise_rhs(st, o)
    names = ["rho0", "rhoinv", "c2rho0", "gradrho0_z", "gradp0_z", "spongexyz", "g"]
    if cfg.mhd:
        names += ["box", "curlbox", "curlboy", "curlboz", "gradrho0_x", "gradrho0_y",
                  "gradp0_x", "gradp0_y", "reduction"]
    for n in names:
        ref = o.array(n)
        np.testing.assert_allclose(st.bg[n], ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max(),
                                   err_msg=n)
    for f, d in st.pml.items():
        for key, on in (("a", "pml_a"), ("b", "pml_b"), ("kappa", "pml_k")):
            ref = o.array(f"{on}{f}")
            np.testing.assert_allclose(d[key].ravel(), ref.ravel(), rtol=1e-13,
                                       atol=1e-13 * np.abs(ref).max(), err_msg=f"{key}{f}")
    np.testing.assert_allclose(st.stretch, o.array("stretch"), rtol=1e-15)
    assert st.deltat == o.scalar("deltat")
    for s in range(5):
        assert st.betas[s] == o.scalar(f"beta{s + 1}")


def test_source_amplitude_and_profiles_against_oracle_forcing():
    from sparc2_amd.initialise import source_profiles
    cfg = quiet_cfg()
    bg = background(cfg)
    o = Oracle(cfg, bg)
    st = prepare(cfg, bg)
    sx, sy, sz = source_profiles(st)
    for t in (3, 17, 40):
        o.forcing(t)
        ref = o.array("forcing")
        got = source_amplitude(st, t) * sz[:, None, None] * sy[None, :, None] * sx[None, None, :]
        np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-300)


def _declared_functions():
    hdr = open(os.path.join(ROOT, "include", "sparc_hip.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(sparc_hip_\w+)\(", hdr, re.M)))


def test_library_exports_every_declared_entry_point():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libsparc_hip.so not built (run __graft_entry__.build())")
    declared = _declared_functions()
    assert len(declared) >= 20
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared:
        assert hasattr(L, name), name
    assert set(declared) == set(_lib.EXPORTS)


def test_h5_library_exports_every_declared_entry_point():
    from sparc2_amd import h5io
    if not os.path.exists(h5io.LIB):
        pytest.skip("libsparc_h5.so not built (no HDF5)")
    hdr = open(os.path.join(ROOT, "include", "sparc_h5.h")).read()
    declared = sorted(set(re.findall(r"^(?:int|const char\*)\s+(sparc_h5_\w+)\(", hdr, re.M)))
    assert len(declared) == 6
    L = ctypes.CDLL(h5io.LIB)
    for name in declared:
        assert hasattr(L, name), name


def test_fits_library_exports_every_declared_entry_point():
    from sparc2_amd import forcing
    if not os.path.exists(forcing.LIB):
        pytest.skip("libsparc_fits.so not built (no CFITSIO)")
    hdr = open(os.path.join(ROOT, "include", "sparc_fits.h")).read()
    declared = sorted(set(re.findall(r"^(?:int|const char\*)\s+(sparc_fits_\w+)\(", hdr, re.M)))
    assert len(declared) == 3
    L = ctypes.CDLL(forcing.LIB)
    for name in declared:
        assert hasattr(L, name), name


def test_library_refuses_to_run_without_a_device():
    """No silent CPU fallback: creating a context without a HIP device fails."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libsparc_hip.so not built")
    if os.path.exists("/dev/kfd"):
        pytest.skip("a ROCm device node is present")
    from sparc2_amd.solver import Solver
    cfg = quiet_cfg()
    with pytest.raises(_lib.SparcError):
        Solver(cfg, background(cfg))
