"""Tile autotuner: tuning leaves the run's state untouched, and the state a
run reaches does not depend on the tile shape (bitwise)."""
import numpy as np
import pytest

from tests.cases import background, mhd_cfg, quiet_cfg, random_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("make", [lambda: quiet_cfg(nx=48, ny=32, nz=40),
                                  lambda: mhd_cfg(nx=48, ny=32, nz=40)], ids=["quiet", "mhd"])
def test_autotune_keeps_state_and_results(make):
    from sparc2_amd.solver import Solver
    cfg = make()
    bg = background(cfg)
    a0, psi0 = random_state(cfg, seed=5)
    ref = Solver(cfg, bg)
    ref.put_state(a0, psi0)
    ref.advance(1, 3)
    want = ref.get_state()

    s = Solver(cfg, bg, setup=ref.st)
    s.put_state(a0, psi0)
    s.advance(1, 1)
    before = s.get_state(with_psi=True)
    tile, ms = s.autotune(reps=2, cache=False)
    after = s.get_state(with_psi=True)
    print("autotuned", tile, ms, "default", ref.tile()[0], flush=True)
    np.testing.assert_array_equal(before[0], after[0])
    for f in before[1]:
        np.testing.assert_array_equal(before[1][f], after[1][f])
    assert s.tile()[0] == tuple(tile) and ms > 0
    s.advance(2, 3)
    np.testing.assert_array_equal(s.get_state(), want)
    # an explicit shape is applied, a shape that does not fit is refused
    s.set_tile((32, 8, 0))
    assert s.tile()[0][:2] == (32, 8)
    from sparc2_amd._lib import SparcError
    with pytest.raises(SparcError):
        s.set_tile((24, 8, 0))
