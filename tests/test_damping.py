"""FFT damping (DAMPING = 1): INIT_DAMPING's rates (initialize.f90:1276-1336)
and damp_velocity (all_modules.f90:1802-1888) in the oracle, pinned by what
the mathematics fixes: a single horizontal Fourier mode is an eigenfunction
of the unnormalised r2c -> x rates -> c2r round trip with eigenvalue
nx*ny*rate(k) (the rates carry the 1/(nx*ny)); a constant field (k = 0, rate
0) gives no term; a random field matches numpy's own FFT (rfft2/irfft2, an
independent implementation); the rates are symmetric in ky."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from sparc2_amd.initialise import damping_rates, prepare
from tests.cases import background, quiet_cfg


def _setup(nx=24, ny=18, nz=16):
    cfg = quiet_cfg(nx=nx, ny=ny, nz=nz, damping=True)
    bg = background(cfg)
    o = Oracle(cfg, bg)
    st = prepare(cfg, bg)
    return cfg, o, st


def test_rates_host_equals_oracle_and_fold_in_ky():
    cfg, o, st = _setup()
    r = damping_rates(st)
    assert r.shape == (cfg.ny, cfg.nx // 2 + 1)
    np.testing.assert_allclose(o.array("damping_rates").reshape(r.shape), r, rtol=1e-14, atol=0)
    assert r[0, 0] == 0.0
    for ky in range(1, cfg.ny):
        np.testing.assert_array_equal(r[ky], r[cfg.ny - ky])
    # grows with |k|
    assert np.all(np.diff(r[0]) > 0) and np.all(np.diff(r[: cfg.ny // 2 + 1, 0]) > 0)


@pytest.mark.parametrize("mode", ["x1", "x5", "y3", "nyquist_x", "diag"])
def test_fourier_modes_are_eigenfunctions(mode):
    cfg, o, st = _setup()
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    r = damping_rates(st)
    x = np.arange(nx)[None, None, :]
    y = np.arange(ny)[None, :, None]
    z = 1.0 + 0.05 * np.arange(nz)[:, None, None]
    if mode == "x1":
        v, rate = np.cos(2 * np.pi * x / nx) * z * np.ones((1, ny, 1)), r[0, 1]
    elif mode == "x5":
        v, rate = np.sin(2 * np.pi * 5 * x / nx) * z * np.ones((1, ny, 1)), r[0, 5]
    elif mode == "y3":
        v, rate = np.cos(2 * np.pi * 3 * y / ny) * z * np.ones((1, 1, nx)), r[3, 0]
    elif mode == "nyquist_x":
        v, rate = ((-1.0) ** x) * z * np.ones((1, ny, 1)), r[0, nx // 2]
    else:
        # cos(kx x + ky y): modes (2, 4) and (-2, -4), rate(2, 4) = rate(2, ny-4)
        v, rate = np.cos(2 * np.pi * (2 * x / nx + 4 * y / ny)) * z, r[4, 2]
    term = o.damp_term(v)
    np.testing.assert_allclose(term, nx * ny * rate * v, rtol=0, atol=1e-12 * nx * ny * rate)


def test_constant_field_is_not_damped():
    """Only rounding survives: the DFT's rounding at k != 0 (~1e-16 of the
    field) times the largest eigenvalue nx*ny*max(rate)."""
    cfg, o, st = _setup()
    term = o.damp_term(np.full((cfg.nz, cfg.ny, cfg.nx), 3.7))
    top = 3.7 * cfg.nx * cfg.ny * damping_rates(st).max()
    assert np.abs(term).max() < 1e-14 * top


def test_random_field_matches_numpy_fft():
    cfg, o, st = _setup(nx=20, ny=16, nz=16)
    rng = np.random.default_rng(4)
    v = rng.standard_normal((cfg.nz, cfg.ny, cfg.nx))
    r = damping_rates(st)
    ref = np.fft.irfft2(np.fft.rfft2(v) * r[None], s=(cfg.ny, cfg.nx)) * (cfg.nx * cfg.ny)
    got = o.damp_term(v)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-13


def test_damped_rhs_subtracts_the_term_after_zeroing():
    """MP_QUIET_3D with DAMPING: scr(v) = undamped RHS - term, also on the
    zeroed boundary planes (RHS.f90:154-172)."""
    cfg, o, st = _setup(nx=16, ny=16, nz=20)
    rng = np.random.default_rng(2)
    a = 1e-4 * rng.standard_normal((5, cfg.nz, cfg.ny, cfg.nx))
    o.put_state(a)
    o.rhs_stage(1)
    damped = o.scr()
    o.set_damping(False)
    o.rhs_stage(1)
    plain = o.scr()
    o.set_damping(True)
    for f in (1, 2, 3):
        t = o.damp_term(a[f])
        assert np.linalg.norm(plain[f] - damped[f] - t) / np.linalg.norm(t) < 1e-12
    np.testing.assert_array_equal(plain[0], damped[0])
    np.testing.assert_array_equal(plain[4], damped[4])
    assert np.abs(damped[3][0]).max() > 0      # boundary plane: zeroed, then damped
