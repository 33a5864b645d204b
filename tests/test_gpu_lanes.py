"""FP32 stage kernels with two x points per lane against the same kernels with
one point per lane (SPARC_LANE_POINTS=1), on the same seeded input.

A two-point lane computes each of its points with the operations, in the
order, of a one-point lane, so the states after N steps agree bitwise; the
diagnostics sum their block partials in another order (relative 1e-12).
The cases cover the x-edge closures, x-face CPML points a lane pair can
straddle (odd face widths: point by point) and even face widths (the pair at
once), for non-periodic quiet with x/y CPML and six-face MHD, periodic wrap,
and an odd nx, which keeps one point per lane.
"""
import os

import numpy as np
import pytest

from tests.cases import background, mhd_cfg, quiet_cfg, random_state

pytestmark = pytest.mark.gpu


def _run(cfg, bg, a, psi, n, one_point, tile):
    from sparc2_amd.solver import Solver
    old = os.environ.get("SPARC_LANE_POINTS")
    os.environ["SPARC_LANE_POINTS"] = "1" if one_point else "2"
    try:
        s = Solver(cfg, bg)
    finally:
        if old is None:
            del os.environ["SPARC_LANE_POINTS"]
        else:
            os.environ["SPARC_LANE_POINTS"] = old
    s.set_tile(tile)
    nblk = s.tile()[1]
    s.put_state(a, psi)
    s.advance(0, n)
    out = s.get_state(with_psi=True)
    d = s.diag()
    s.close()
    return out, d, nblk


CASES = {
    "quiet_periodic": lambda: quiet_cfg(nx=64, ny=40, nz=48, precision="fp32"),
    # x edges: one-sided rows, zeroed edge planes and 5-point x CPML faces
    # (odd width: the face edge falls inside a lane pair)
    "quiet_hpml": lambda: quiet_cfg(nx=56, ny=40, nz=40, periodic=False, use_hpml=True,
                                    npml=(5, 5, 5, 5, 8, 8), precision="fp32"),
    "mhd_six_faces": lambda: mhd_cfg(nx=48, ny=40, nz=40, npml=(7, 7, 6, 6, 7, 7),
                                     precision="fp32"),
    # even x-face widths: lane pairs go through the x-face CPML together
    "quiet_hpml_even_x": lambda: quiet_cfg(nx=56, ny=40, nz=40, periodic=False, use_hpml=True,
                                           npml=(6, 6, 5, 5, 8, 8), precision="fp32"),
    "mhd_even_x": lambda: mhd_cfg(nx=48, ny=40, nz=40, npml=(10, 10, 7, 7, 7, 7),
                                  precision="fp32"),
    "mhd_periodic": lambda: mhd_cfg(nx=64, ny=32, nz=40, periodic=True, precision="fp32"),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("tile", [(32, 8, 0), (16, 16, 0)])
def test_two_point_lanes_match_one_point_lanes(case, tile):
    cfg = CASES[case]()
    bg = background(cfg)
    a, psi = random_state(cfg, seed=11)
    print(f"lanes {case} tile {tile}", flush=True)
    (s2, psi2), d2, nb2 = _run(cfg, bg, a, psi, 12, False, tile)
    (s1, psi1), d1, nb1 = _run(cfg, bg, a, psi, 12, True, tile)
    # two points per lane: half the tiles across x
    assert nb2 < nb1
    assert np.abs(s1[3]).max() > 0
    np.testing.assert_array_equal(s2, s1)
    for f in psi1:
        np.testing.assert_array_equal(psi2[f], psi1[f])
    np.testing.assert_allclose(d2, d1, rtol=1e-12, atol=0)


def test_odd_nx_keeps_one_point_lanes():
    cfg = quiet_cfg(nx=45, ny=32, nz=40, precision="fp32")
    bg = background(cfg)
    a, psi = random_state(cfg, seed=5)
    (s2, _), _, nb2 = _run(cfg, bg, a, psi, 6, False, (32, 8, 0))
    (s1, _), _, nb1 = _run(cfg, bg, a, psi, 6, True, (32, 8, 0))
    assert nb2 == nb1
    np.testing.assert_array_equal(s2, s1)
