"""Precision form: FP32 fields/stencils/CPML with the compensated (Kahan) RK
update, against the plain FP64 GPU path on the same input.

Tolerance: relative L2 <= 1e-5 per field (the bar stated for the
FP32-compensated forms versus the FP64 oracle).
"""
import numpy as np
import pytest

from tests.cases import background, field_errors, mhd_cfg, quiet_cfg

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _run(cfg, bg, n):
    from sparc2_amd.solver import Solver
    s = Solver(cfg, bg)
    s.advance(0, n)
    a = s.get_state()
    s.close()
    return a


@pytest.mark.parametrize("case,n", [("quiet", 300), ("mhd", 200)])
def test_fp32_compensated_tracks_fp64(case, n):
    cfg64 = quiet_cfg(nx=48, ny=48, nz=64) if case == "quiet" else mhd_cfg(nx=48, ny=48, nz=48)
    bg = background(cfg64)
    cfg32 = cfg64.replace(precision="fp32")
    ref = _run(cfg64, bg, n)
    got = _run(cfg32, bg, n)
    assert np.abs(ref[3]).max() > 0
    errs = field_errors(got, ref)
    print(f"{case} fp32 vs fp64 after {n + 1} steps: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) < TOL
