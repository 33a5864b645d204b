"""The compiled Fortran drop-in shim (sparc2_amd/fortran/sparc_gpu.f90):
a Fortran program fills the reference's module globals (the stand-in
all_modules of tests/fortran), hands over both z CPML faces the way a
per-rank INITIALISE_RHS leaves them, and runs the time loop of
driver.f90:123-202 with the shim's argument-less STEP and FILTER_VARS and
its device-side failure checks.  Its final state and CPML bases equal the
oracle's from the same inputs within relative L2 1e-11."""
import os
import subprocess

import numpy as np
import pytest

from oracle.oracle import Oracle
from tests.cases import background, field_errors, quiet_cfg, random_state, rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "_build", "drive_quiet")


def _write_input(path, cfg, st, a, psi, nmax):
    hi = np.zeros(24, dtype=np.int32)
    hi[:15] = [cfg.nx, cfg.ny, cfg.nz, 1, int(cfg.periodic), int(cfg.use_pml), cfg.npml[4],
               cfg.npml[5], cfg.g_xy_filt, cfg.g_z_filt, cfg.xyfreq, cfg.zfreq, int(cfg.pulse),
               cfg.pulse_dir, nmax]
    hr = np.zeros(32)
    hr[:25] = [cfg.s_xy, cfg.s_z, st.deltat, *st.betas, st.stretchx, st.stretchy, cfg.timestep,
               cfg.pulse_amp, cfg.pulse_t0, cfg.pulse_t1, cfg.pulse_st, cfg.pulse_px, cfg.pulse_py,
               cfg.pulse_pz, cfg.pulse_sr, cfg.pulse_sz, st.dimc, st.diml, st.dimrho, cfg.xlength,
               cfg.ylength]
    with open(path, "wb") as fh:
        fh.write(hi.tobytes())
        fh.write(hr.tobytes())
        for v in (st.xaxis, st.yaxis, st.zaxis, st.stretch, st.bg["g"]):
            fh.write(np.ascontiguousarray(v, dtype=np.float64).tobytes())
        for n in ("rho0", "rhoinv", "c2rho0", "gradrho0_z", "gradp0_z", "spongexyz", "p0"):
            v = np.broadcast_to(np.asarray(st.bg[n], dtype=np.float64), (cfg.nz, cfg.ny, cfg.nx))
            fh.write(np.ascontiguousarray(v).tobytes())
        fh.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        for f in (4, 5):
            if cfg.face_active(f):
                P = st.pml[f]
                for key in ("a", "b", "kappa"):
                    fh.write(np.ascontiguousarray(P[key], dtype=np.float64).tobytes())
                fh.write(np.ascontiguousarray(psi[f], dtype=np.float64).tobytes())


@pytest.mark.skipif(not os.path.exists(EXE), reason="shim test program not built (needs flang)")
def test_fortran_program_with_the_drop_in_step_matches_the_oracle(tmp_path):
    from sparc2_amd.solver import Solver
    cfg = quiet_cfg(nx=32, ny=32, nz=48)
    assert cfg.face_active(4) and cfg.face_active(5) and cfg.pulse
    bg = background(cfg)
    s = Solver(cfg, bg)                  # the product's FP64 set-up (INITIALISE_RHS)
    st = s.st
    s.close()
    a, psi = random_state(cfg, seed=51)
    nmax = 2 * cfg.xyfreq * cfg.zfreq + 1     # covers both filter cadences
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(fin, cfg, st, a, psi, nmax)
    print(f"running {EXE} for {nmax + 1} steps", flush=True)
    r = subprocess.run([EXE, fin, fout], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    raw = np.fromfile(fout, dtype=np.float64)
    nst = cfg.nv * cfg.npoints
    got = raw[:nst].reshape((cfg.nv, cfg.nz, cfg.ny, cfg.nx))
    o = Oracle(cfg, bg)
    o.put_state(a)
    for f, v in psi.items():
        o.put_psi(f, v)
    o.run(0, nmax)
    errs = field_errors(got, o.get_state())
    print("Fortran shim vs oracle: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) < 1e-11
    off = nst
    for f in (4, 5):
        n = cfg.npsi * int(np.prod(cfg.face_shape(f)))
        e = rel_l2(raw[off:off + n], o.get_psi(f))
        print(f"psi face {f}: {e:.2e}", flush=True)
        assert e < 1e-11
        off += n
    assert off == raw.size
