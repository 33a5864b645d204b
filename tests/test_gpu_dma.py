"""The MHD xy kernel's two staging paths give the same bits.

On 16x16 tiles whose rows hold an even number of 8-byte lanes (nx even in
FP64, a multiple of 4 in FP32 with two points per lane) the kernel stages
its planes with direct-to-LDS buffer loads (two lane cells per lane, packed
cross-shaped LDS region, fluxes formed in LDS); SPARC_MHD_DMA=0 (read when a
context is made) keeps the register staging.  Both read the same values and
do the same arithmetic, so states and CPML bases after a driver loop with
filters must be identical, on a 6-face CPML grid, a periodic one, and grids
the direct path refuses (odd nx; in FP32 nx = 2 mod 4), where the register
path runs either way.
"""
import os

import numpy as np
import pytest

from tests.cases import background, mhd_cfg, random_state

pytestmark = pytest.mark.gpu

CASES = {
    "six_faces": lambda: mhd_cfg(nx=48, ny=38, nz=40),
    "periodic": lambda: mhd_cfg(nx=64, ny=32, nz=36, periodic=True),
    "odd_nx": lambda: mhd_cfg(nx=47, ny=38, nz=36),
    "nx_2_mod_4": lambda: mhd_cfg(nx=54, ny=38, nz=36),
}


def _run(cfg, bg, a, psi, dma, nsteps=9):
    from sparc2_amd.solver import Solver
    old = os.environ.get("SPARC_MHD_DMA")
    os.environ["SPARC_MHD_DMA"] = "1" if dma else "0"
    try:
        s = Solver(cfg, bg, tile=(16, 16, 0))
    finally:
        if old is None:
            del os.environ["SPARC_MHD_DMA"]
        else:
            os.environ["SPARC_MHD_DMA"] = old
    s.put_state(a, psi)
    s.advance(0, nsteps - 1)
    out = s.get_state(with_psi=True)
    s.close()
    return out


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_direct_to_lds_staging_equals_register_staging(case, precision):
    cfg = CASES[case]().replace(precision=precision)
    bg = background(cfg)
    a, psi = random_state(cfg, seed=29)
    assert cfg.xyfreq <= 9 and cfg.zfreq <= 9     # filters inside the loop
    got, gpsi = _run(cfg, bg, a, psi, dma=True)
    ref, rpsi = _run(cfg, bg, a, psi, dma=False)
    print(f"{case} {precision} ({cfg.nx}x{cfg.ny}x{cfg.nz}): max |diff| {np.abs(got - ref).max():.3e}",
          flush=True)
    np.testing.assert_array_equal(got, ref)
    for f in rpsi:
        np.testing.assert_array_equal(gpsi[f], rpsi[f])
