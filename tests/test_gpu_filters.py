"""FILTER_VARS alone (derivatives.f90:659-703) on the GPU against the oracle.

The y and x passes run as one kernel: a wave owns 64 consecutive x points
and writes the inner 52 (g = 5) or 56 (g = 3) of them, reading the x stencil
from the neighbouring lanes; it marches up chunks of 64 rows.  The grids
here put nx below, at and just above one wave's output width and ny across
two row chunks, periodic and not, with both filter widths, so every lane
mask, wrap and chunk seam is exercised (non-periodic grids keep the stage
kernels' rule of at least 6 cells in a boundary tile).  FP64 must equal the
oracle to 1e-13 (the same arithmetic in the same order); the FP32 form
(compensation terms stored as bfloat16 and filtered alike) stays within 1e-6
of FP64.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from tests.cases import background, field_errors, mhd_cfg, quiet_cfg, random_state

pytestmark = pytest.mark.gpu

G3 = dict(filter_mode_xy=2, filter_mode_z=2, s_xy=0.7, s_z=0.6)

CASES = {
    "quiet_periodic_nx20": lambda: quiet_cfg(nx=20, ny=24, nz=20),
    "quiet_periodic_nx53_ny70": lambda: quiet_cfg(nx=53, ny=70, nz=18),
    "quiet_periodic_g3_nx56": lambda: quiet_cfg(nx=56, ny=20, nz=18, **G3),
    "quiet_hpml_nx52": lambda: quiet_cfg(nx=52, ny=40, nz=20, use_hpml=True,
                                         npml=(6, 6, 6, 6, 8, 8)),
    "quiet_hpml_g3_nx70_ny70": lambda: quiet_cfg(nx=70, ny=70, nz=18, use_hpml=True,
                                                 npml=(6, 6, 6, 6, 8, 8), **G3),
    "mhd_nx38": lambda: mhd_cfg(nx=38, ny=22, nz=20),
    "mhd_periodic_nx116": lambda: mhd_cfg(nx=116, ny=18, nz=18, periodic=True),
}


def _filter_time(cfg):
    """A time index on both filter cadences."""
    return cfg.xyfreq * cfg.zfreq


@pytest.mark.parametrize("case", sorted(CASES))
def test_filter_vars_matches_oracle(case):
    from sparc2_amd.solver import Solver
    cfg = CASES[case]()
    bg = background(cfg)
    a, _ = random_state(cfg, seed=17)
    t = _filter_time(cfg)
    o = Oracle(cfg, bg)
    o.put_state(a)
    o.filter_vars(t)
    ref = o.get_state()
    s = Solver(cfg, bg)
    s.put_state(a)
    s.filter_vars(t)
    got = s.get_state()
    errs = field_errors(got, ref)
    print(f"{case} ({cfg.nx}x{cfg.ny}x{cfg.nz}, g = {cfg.g_xy_filt}/{cfg.g_z_filt}, "
          f"periodic {cfg.periodic}): " + " ".join(f"{e:.1e}" for e in errs), flush=True)
    assert max(errs) < 1e-13
    # the filters changed the state (the comparison is not of two copies)
    assert np.abs(ref - a).max() > 1e-3 * np.abs(a).max()
    s32 = Solver(cfg.replace(precision="fp32"), bg, setup=s.st)
    s32.put_state(a)
    s32.filter_vars(t)
    e32 = field_errors(s32.get_state(), got)
    print(f"  FP32 vs FP64: " + " ".join(f"{e:.1e}" for e in e32), flush=True)
    assert max(e32) < 1e-6
    s32.close()
    s.close()
