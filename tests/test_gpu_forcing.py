"""PULSE = 0 driver on the GPU: a forcing cube read from a FITS file,
LAGRANGE_INTERP x source_damp per step, times source_dep added to RHSv_z in
the stage kernels (RHS.f90:146-147, 640-641), against the oracle fed the same
per-step maps.  Bar: FP64 relative L2 1e-11 per field after N driver steps."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from sparc2_amd import forcing as fz
from tests.cases import background, field_errors, mhd_cfg, quiet_cfg, random_state

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not __import__("os").path.exists(fz.LIB), reason="libsparc_fits.so not built")]


def _cube(cfg, n, seed):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((n, cfg.ny, cfg.nx))
    for ax in (1, 2):
        c = 0.25 * np.roll(c, 1, ax) + 0.5 * c + 0.25 * np.roll(c, -1, ax)
    return c


@pytest.mark.parametrize("case", ["quiet", "mhd"])
def test_fits_forcing_driver_matches_oracle(tmp_path, case):
    from sparc2_amd.solver import Solver
    base = dict(pulse=False, solartime=0.01, cadforcing=2.0, randomexcitedepth=-0.2e8,
                randomexcitesigma=0.3e8, randomexciteampl=1.0e4)
    cfg = quiet_cfg(**base) if case == "quiet" else mhd_cfg(**base)
    bg = background(cfg)
    cube = _cube(cfg, fz.num_steps(cfg), 3)
    path = str(tmp_path / "forcing.fits")
    fz.write_fits_cube(path, cube)
    cfg = cfg.replace(forcingfunc=path)
    s = Solver(cfg, bg)
    assert s.series is not None
    a, psi = random_state(cfg, seed=4, amp=1e-6)
    s.put_state(a, psi)
    o = Oracle(cfg, bg)
    o.put_state(a)
    for f, v in psi.items():
        o.put_psi(f, v)
    series = fz.ForcingSeries(s.st, cube)
    sdep = fz.source_dep(s.st)
    nsteps = 14
    print(f"forcing {case}: {nsteps} steps", flush=True)
    for t in range(nsteps + 1):
        o.set_forcing_map(series.forcing(t), sdep if t == 0 else None)
        o.step()
        if t != 0:
            o.filter_vars(t)
    s.advance(0, nsteps)
    got, ref = s.get_state(), o.get_state()
    errs = field_errors(got, ref)
    print("errors: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) < 1e-11
    # the forcing acts: the same run without it ends far from this one
    s.close()
    plain = Solver(cfg.replace(forcingfunc=""), bg)
    plain.put_state(a, psi)
    plain.advance(0, nsteps)
    assert field_errors(plain.get_state(), got)[3] > 1e3 * 1e-11
    plain.close()
