"""FFT damping on the GPU (hipFFT r2c/c2r per z plane, the rates multiply,
the term subtracted after the boundary zeroing) against the oracle's direct
DFT sums, through the C ABI.  Bar: FP64 relative L2 1e-11 per field after N
driver steps; FP32 (two-point lanes, Kahan update) within 1e-5 of FP64."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from tests.cases import background, field_errors, mhd_cfg, quiet_cfg, random_state

pytestmark = pytest.mark.gpu

TOL_STEP = 1e-11

CASES = {
    "quiet": lambda **kw: quiet_cfg(damping=True, **kw),
    "quiet_hpml": lambda **kw: quiet_cfg(damping=True, periodic=False, use_hpml=True,
                                         npml=(5, 6, 6, 5, 8, 7), **kw),
    "mhd": lambda **kw: mhd_cfg(damping=True, **kw),
}


def _run_gpu(cfg, bg, a, psi, n):
    from sparc2_amd.solver import Solver
    s = Solver(cfg, bg)
    s.put_state(a, psi)
    s.advance(0, n)
    out = s.get_state()
    s.close()
    return out


@pytest.mark.parametrize("case,nsteps", [("quiet", 12), ("quiet_hpml", 8), ("mhd", 8)])
def test_damped_driver_loop_matches_oracle(case, nsteps):
    cfg = CASES[case]()
    bg = background(cfg)
    a, psi = random_state(cfg, seed=21)
    o = Oracle(cfg, bg)
    o.put_state(a)
    for f, v in psi.items():
        o.put_psi(f, v)
    o.run(0, nsteps)
    ref = o.get_state()
    print(f"damping {case} {nsteps} steps", flush=True)
    got = _run_gpu(cfg, bg, a, psi, nsteps)
    errs = field_errors(got, ref)
    print("errors: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) < TOL_STEP
    # the damping acts: the undamped run differs by far more than the bar
    plain = _run_gpu(cfg.replace(damping=False), bg, a, psi, nsteps)
    assert max(field_errors(plain, ref)[1:4]) > 1e3 * TOL_STEP


def test_damped_fp32_tracks_fp64():
    cfg = CASES["quiet"](nx=48, ny=48, nz=48)
    bg = background(cfg)
    a, psi = random_state(cfg, seed=22)
    ref = _run_gpu(cfg, bg, a, psi, 60)
    got = _run_gpu(cfg.replace(precision="fp32"), bg, a, psi, 60)
    errs = field_errors(got, ref)
    print("damping fp32 vs fp64: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) < 1e-5


def test_damped_runs_are_not_batched():
    from sparc2_amd._lib import SparcError
    from sparc2_amd.batch import Batch
    from sparc2_amd.solver import Solver
    cfg = CASES["quiet"]()
    s = Solver(cfg, background(cfg))
    with pytest.raises(SparcError):
        Batch([s], use_graph=False)
    s.close()
