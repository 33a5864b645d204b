"""Restart through the reference's snapshot layout: a run written to
full_NNNN.h5 at step n and read into a fresh context continues as the
uninterrupted run does.  The file holds physical units (a*dimc and so on,
all_modules.f90:1137-1185), so a value comes back within one rounding of the
scaling; the continued state and CPML bases then agree to 1e-13 relative
L2 per field and face."""
import os

import pytest

from sparc2_amd import h5io
from tests.cases import background, field_errors, mhd_cfg, quiet_cfg, random_state, rel_l2

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(h5io.LIB), reason="libsparc_h5.so not built")]


@pytest.mark.parametrize("case", ["quiet", "mhd"])
def test_restart_from_full_snapshot(tmp_path, case):
    from sparc2_amd.solver import Solver
    cfg = quiet_cfg() if case == "quiet" else mhd_cfg()
    bg = background(cfg)
    a, psi = random_state(cfg, seed=8)
    s = Solver(cfg, bg)
    s.put_state(a, psi)
    s.advance(0, 6)
    p = str(tmp_path / h5io.file_name("run_", "full", 6))
    h5io.write_full(p, s, time=6)
    s.advance(7, 14)
    ref, ref_psi = s.get_state(with_psi=True)
    r = Solver(cfg, bg, setup=s.st)
    assert h5io.read_full(p, r) == pytest.approx(6 * cfg.timestep / 3600.0)
    r.advance(7, 14)
    got, got_psi = r.get_state(with_psi=True)
    errs = field_errors(got, ref)
    print(f"restart {case}: " + " ".join(f"{e:.1e}" for e in errs), flush=True)
    assert max(errs) < 1e-13
    for f in ref_psi:
        assert rel_l2(got_psi[f], ref_psi[f]) < 1e-13, f
    s.close()
    r.close()
