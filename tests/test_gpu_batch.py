"""Batched form: runs of different sizes and physics advanced together give
the same state as each run advanced alone (bit for bit: the batched launch
runs the same per-tile code on the same inputs)."""
import numpy as np
import pytest

from tests.cases import background, mhd_cfg, quiet_cfg, random_state

pytestmark = pytest.mark.gpu


def _cases(precision):
    cfgs = [
        quiet_cfg(nx=32, ny=32, nz=40, precision=precision),
        quiet_cfg(nx=48, ny=32, nz=56, pulse_px=20e8, precision=precision),
        quiet_cfg(nx=32, ny=48, nz=40, periodic=False, use_hpml=True, npml=(5, 6, 6, 5, 8, 7),
                  precision=precision),
        mhd_cfg(nx=32, ny=32, nz=40, precision=precision),
        mhd_cfg(nx=48, ny=32, nz=40, periodic=True, precision=precision),
    ]
    out = []
    for q, cfg in enumerate(cfgs):
        bg = background(cfg)
        a, psi = random_state(cfg, seed=20 + q)
        out.append((cfg, bg, a, psi))
    return out


def _solvers(cases):
    from sparc2_amd.solver import Solver
    ss = []
    for cfg, bg, a, psi in cases:
        s = Solver(cfg, bg)
        s.put_state(a, psi)
        ss.append(s)
    return ss


@pytest.mark.parametrize("precision,use_graph,autotune", [("fp64", False, False), ("fp64", True, False),
                                                         ("fp32", True, False), ("fp32", True, True)])
def test_batch_equals_alone(precision, use_graph, autotune):
    from sparc2_amd.batch import Batch
    cases = _cases(precision)
    nsteps = 13
    alone = []
    for s in _solvers(cases):
        s.advance(0, nsteps - 1)
        alone.append(s.get_state(with_psi=True))
        s.close()
    ss = _solvers(cases)
    b = Batch(ss, use_graph=use_graph, autotune=autotune)
    print("groups, blocks:", b.info(), "tiles:", b.tiles, flush=True)
    b.advance(5)          # two calls: the second continues the time index
    b.advance(nsteps - 5)
    for q, s in enumerate(ss):
        a, psi = s.get_state(with_psi=True)
        ra, rpsi = alone[q]
        print(q, np.abs(a - ra).max(), flush=True)
        np.testing.assert_array_equal(a, ra)
        for f in rpsi:
            np.testing.assert_array_equal(psi[f], rpsi[f])
        d = s.diag()
        assert np.all(np.isfinite(d))
    b.close()
