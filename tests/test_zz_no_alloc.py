"""The allocation-free derivative (sparc_hip_deriv_noalloc) against the
allocating one (sparc_hip_ddx/ddy/ddz), through the C ABI: bit for bit (the
same kernel on the same input), on two calls in a row on one device, the
second on a smaller grid; a grid above the staging limit is refused."""
import ctypes as C

import numpy as np
import pytest

from tests.cases import background, quiet_cfg

pytestmark = pytest.mark.gpu

AXES = (("ddx", 0), ("ddy", 1), ("ddz", 2))


def _noalloc(s, axis, x, bc):
    from sparc2_amd._lib import check, dptr, lib
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.full(s.shape, np.nan)
    check(lib().sparc_hip_deriv_noalloc(s.h, int(axis), dptr(x), dptr(out), int(bc)),
          "sparc_hip_deriv_noalloc")
    return out


def test_noalloc_derivative_equals_the_allocating_one():
    from sparc2_amd.solver import Solver
    rng = np.random.default_rng(61)
    for cfg in (quiet_cfg(nx=48, ny=40, nz=56, periodic=False, use_hpml=True,
                          npml=(5, 5, 5, 5, 8, 8)),
                quiet_cfg(nx=32, ny=32, nz=40)):           # second call: a smaller grid
        s = Solver(cfg, background(cfg))
        x = rng.standard_normal(s.shape)
        bc = 5 if cfg.periodic else 1
        for name, axis in AXES:
            b = 1 if axis == 2 else bc
            want = getattr(s, name)(x, b)
            got = _noalloc(s, axis, x, b)
            print(cfg.nx, cfg.ny, cfg.nz, name, np.abs(got - want).max(), flush=True)
            np.testing.assert_array_equal(got, want)
        s.close()


def test_noalloc_derivative_refuses_large_grids():
    from sparc2_amd._lib import SparcError
    from sparc2_amd.solver import Solver
    cfg = quiet_cfg(nx=256, ny=256, nz=72)                  # 4718592 > 4194304 points
    s = Solver(cfg, background(cfg))
    with pytest.raises(SparcError, match="exceeds"):
        _noalloc(s, 0, np.zeros(s.shape), 5)
    s.close()
