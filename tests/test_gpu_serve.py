"""Served form: requests arriving while others run join the batch; each
result (photospheric vz slices and final state) is the same as the request
served alone (bitwise: the batched launch runs the same per-tile code)."""
import copy
import os

import numpy as np
import pytest

from tests.cases import mhd_cfg, quiet_cfg

pytestmark = pytest.mark.gpu


def _requests(precision="fp32"):
    from sparc2_amd.serve import Request
    specs = [
        (0.00, quiet_cfg(nx=32, ny=32, nz=40, precision=precision), (-10.0, 0.5), 140, 0.0),
        (0.02, mhd_cfg(nx=32, ny=32, nz=40, precision=precision), (-5.0, 1.0), 90, 2500.0),
        (0.04, quiet_cfg(nx=48, ny=32, nz=48, pulse_px=20e8, precision=precision), (-10.0, 0.5), 60, 0.0),
        (0.30, mhd_cfg(nx=48, ny=32, nz=40, periodic=True, precision=precision), (-5.0, 1.0), 75, 1500.0),
    ]
    return [Request(rid=q, arrival=t, cfg=c, zrange=z, nsteps=n, B0=b, slice_every_s=15.0)
            for q, (t, c, z, n, b) in enumerate(specs)]


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_served_results_equal_alone(tmp_path, precision):
    from sparc2_amd import h5io
    from sparc2_amd.serve import Server, cfl_timestep
    reqs = _requests(precision)
    chunks = str(tmp_path) if os.path.exists(h5io.LIB) else None
    mixed = Server(max_chunk=25, keep_final=True, chunk_dir=chunks).run(copy.deepcopy(reqs), verbose=True)
    for r in reqs:
        alone = Server(max_chunk=25, keep_final=True).run([copy.deepcopy(r)])
        a = mixed["results"][r.rid]
        b = alone["results"][r.rid]
        assert a.steps == b.steps == r.nsteps
        assert len(a.slices) == len(b.slices) > 0
        for (ta, sa), (tb, sb) in zip(a.slices, b.slices):
            assert ta == tb
            np.testing.assert_array_equal(sa, sb)
        np.testing.assert_array_equal(a.final, b.final)
        if chunks:
            # each slice also went out as a chunk file of the surface plane:
            # vz in physical units (x dimc, the sound speed at the first
            # point), time in hours of the run's CFL time step
            bg = r.build_background()
            dimc = float(bg["cs0"][0, 0, 0])
            dt = cfl_timestep(r.cfg, bg)
            for n, (t, sl) in enumerate(a.slices, start=1):
                path = os.path.join(chunks, h5io.file_name(f"req{r.rid}_", "chunk", n))
                dims, hours = h5io.attrs(path)
                assert dims == (r.cfg.nx, r.cfg.ny, 1)
                assert hours == pytest.approx(t * dt / 3600.0, rel=1e-14)
                vz = h5io.read(path, "vz", (1, r.cfg.ny, r.cfg.nx))[0]
                np.testing.assert_allclose(vz, sl * dimc, rtol=1e-15, atol=0)
    print({k: v for k, v in mixed.items() if k != "results"}, flush=True)
    assert mixed["latency_p99"] >= mixed["latency_p50"] > 0
