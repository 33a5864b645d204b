"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no tests, fixtures or golden data (its ./TEST_CASE/ is
absent), so each oracle function is checked against closed forms, invariants
or brute force chosen so that a plausible transcription mistake (a wrong
coefficient, sign, index or edge rule) fails at least one check.
"""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle, coeffs
from tests.cases import background, mhd_cfg, quiet_cfg

DP, F11, F7 = coeffs()
A = DP[5, 6:]          # a_1..a_5 of the centred row


# --------------------------------------------------------------- coefficients
def test_centred_row_is_antisymmetric_and_first_order_exact():
    assert DP[5, 5] == 0.0
    np.testing.assert_array_equal(DP[5, :5], -DP[5, 6:][::-1])
    j = np.arange(1, 6)
    assert abs(2 * np.sum(A * j) - 1.0) < 1e-11         # derivative of a line
    assert abs(2 * np.sum(A * j ** 3)) < 1e-10           # 4th-order consistency


def test_centred_row_modified_wavenumber():
    # DRP optimisation of Bogey & Bailly 2004: k*dx at theta = pi/2
    kstar = 2 * (A[0] - A[2] + A[4])
    assert abs(kstar - 1.56984317474) < 1e-10


@pytest.mark.parametrize("r", range(5))
def test_one_sided_rows_exact_to_second_order(r):
    m = np.arange(11)
    c = DP[r]
    assert abs(np.sum(c)) < 2e-11
    assert abs(np.sum(c * (m - r)) - 1.0) < 2e-11
    assert abs(np.sum(c * (m - r) ** 2)) < 5e-10


def test_filter_transfer_functions():
    th = np.linspace(0, np.pi, 33)
    m11 = np.arange(-5, 6)
    m7 = np.arange(-3, 4)
    h11 = np.array([np.sum(F11 * np.cos(t * m11)) for t in th])
    h7 = np.array([np.sum(F7 * np.cos(t * m7)) for t in th])
    np.testing.assert_allclose(h11, 1.0 - np.sin(th / 2) ** 10, atol=1e-14)
    np.testing.assert_allclose(h7, np.sin(th / 2) ** 6, atol=1e-14)


def test_rk_weights_reproduce_the_lddrk_polynomial():
    cfg = quiet_cfg()
    o = Oracle(cfg, background(cfg))
    dt = o.scalar("deltat")
    b = [o.scalar(f"beta{s}") / dt for s in range(1, 6)]
    assert b[0] == 0.0
    prods = [b[4], b[4] * b[3], b[4] * b[3] * b[2], b[4] * b[3] * b[2] * b[1]]
    np.testing.assert_allclose(prods, [0.5, 0.1665579725791151184, 0.03950410250586729987,
                                       0.007810706393337838236], rtol=1e-15)


# ---------------------------------------------------------------- derivatives
@pytest.fixture(scope="module")
def quiet_oracle():
    cfg = quiet_cfg()
    return cfg, Oracle(cfg, background(cfg))


@pytest.fixture(scope="module")
def open_oracle():
    cfg = quiet_cfg(periodic=False, use_hpml=True, npml=(5, 5, 5, 5, 8, 8))
    return cfg, Oracle(cfg, background(cfg))


@pytest.mark.parametrize("mode", [1, 3, 7])
def test_periodic_ddx_ddy_of_a_sinusoid(quiet_oracle, mode):
    cfg, o = quiet_oracle
    i = np.arange(cfg.nx)
    th = 2 * np.pi * mode / cfg.nx
    f = np.broadcast_to(np.sin(th * i)[None, None, :], (cfg.nz, cfg.ny, cfg.nx))
    kstar = 2 * np.sum(A * np.sin(np.arange(1, 6) * th))
    expect = o.scalar("stretchx") * kstar * np.cos(th * i)
    got = o.ddx(f, 5)
    np.testing.assert_allclose(got, np.broadcast_to(expect, got.shape), atol=1e-12 * o.scalar("stretchx"))
    fy = np.transpose(f, (0, 2, 1))
    goty = o.ddy(np.ascontiguousarray(fy), 5)
    expy = o.scalar("stretchy") * kstar * np.cos(th * i)
    np.testing.assert_allclose(goty, np.broadcast_to(expy[None, :, None], goty.shape),
                               atol=1e-12 * o.scalar("stretchy"))


def test_nonperiodic_ddx_ddy_exact_for_lines_and_parabolas(open_oracle):
    cfg, o = open_oracle
    x = np.arange(cfg.nx, dtype=float)
    lin = np.broadcast_to(3.0 * x[None, None, :], (cfg.nz, cfg.ny, cfg.nx))
    np.testing.assert_allclose(o.ddx(lin, 1), 3.0 * o.scalar("stretchx"), rtol=1e-10)
    par = np.broadcast_to((x * x)[None, None, :], (cfg.nz, cfg.ny, cfg.nx))
    got = o.ddx(par, 1)[0, 0]
    np.testing.assert_allclose(got, 2 * x * o.scalar("stretchx"), rtol=1e-9,
                               atol=1e-9 * o.scalar("stretchx"))
    y = np.arange(cfg.ny, dtype=float)
    liny = np.broadcast_to(-2.0 * y[None, :, None], (cfg.nz, cfg.ny, cfg.nx))
    np.testing.assert_allclose(o.ddy(liny, 1), -2.0 * o.scalar("stretchy"), rtol=1e-10)


def test_ddz_of_the_axis_is_one_except_the_zero_ghost_bottom(quiet_oracle):
    """stretch = 1/(dz/dk) uses the one-sided rows at the bottom, while ddz
    there uses the centred row over zero ghosts (derivatives.f90:509-518 vs
    initialize.f90:676-678): d(z)/dz = 1 everywhere except the 5 bottom rows."""
    cfg, o = quiet_oracle
    z = o.array("zaxis")
    f = np.ascontiguousarray(np.broadcast_to(z[:, None, None], (cfg.nz, cfg.ny, cfg.nx)))
    d = o.ddz(f, 1)[:, 0, 0]
    np.testing.assert_allclose(d[5:], 1.0, rtol=1e-10)
    assert np.all(np.abs(d[:5] - 1.0) > 1e-3)


def test_ddz_of_a_constant_shows_the_zero_ghost_signature(quiet_oracle):
    cfg, o = quiet_oracle
    st = o.array("stretch")
    d = o.ddz(np.full((cfg.nz, cfg.ny, cfg.nx), 2.0), 1)[:, 0, 0]
    np.testing.assert_allclose(d[5:], 0.0, atol=1e-9 * st.max())
    # bottom row k: the missing ghosts k-5..-1 carry coefficients DP[5, 0..4-k]
    for k in range(5):
        expect = st[k] * 2.0 * np.sum(DP[5, 5 - k:])
        assert abs(d[k] - expect) < 1e-9 * st[k]
    with pytest.raises(ValueError):
        o.ddz(np.zeros((cfg.nz, cfg.ny, cfg.nx)), 5)


def test_filters_dc_nyquist_and_edges(quiet_oracle, open_oracle):
    cfg, o = quiet_oracle
    shape = (cfg.nz, cfg.ny, cfg.nx)
    c = np.full(shape, 1.5)
    np.testing.assert_allclose(o.filtx(c), 1.5, rtol=1e-15)
    i = np.arange(cfg.nx)
    nyq = np.ascontiguousarray(np.broadcast_to(((-1.0) ** i)[None, None, :], shape))
    np.testing.assert_allclose(o.filtx(nyq), 0.0, atol=1e-15)
    # z: the g layers at each end are copied (derivatives.f90:624,635)
    rng = np.random.default_rng(0)
    r = rng.standard_normal(shape)
    fz = o.filtz(r)
    np.testing.assert_array_equal(fz[:5], r[:5])
    np.testing.assert_array_equal(fz[-5:], r[-5:])
    cfg2, o2 = open_oracle
    r2 = rng.standard_normal((cfg2.nz, cfg2.ny, cfg2.nx))
    fx = o2.filtx(r2)
    np.testing.assert_array_equal(fx[:, :, :5], r2[:, :, :5])
    np.testing.assert_array_equal(fx[:, :, -5:], r2[:, :, -5:])


def test_seven_point_filter_zeroes_open_edges():
    cfg = quiet_cfg(periodic=False, use_hpml=True, npml=(5, 5, 5, 5, 8, 8), filter_mode_xy=2)
    o = Oracle(cfg, background(cfg))
    rng = np.random.default_rng(1)
    r = rng.standard_normal((cfg.nz, cfg.ny, cfg.nx))
    fy = o.filty(r)
    assert np.all(fy[:, :3, :] == 0.0) and np.all(fy[:, -3:, :] == 0.0)
    i = np.arange(cfg.nx)
    nyq = np.ascontiguousarray(np.broadcast_to(((-1.0) ** i)[None, None, :], r.shape))
    np.testing.assert_allclose(o.filtx(nyq)[:, :, 3:-3], nyq[:, :, 3:-3], rtol=1e-14)


# --------------------------------------------------------------------- RK step
def test_step_amplification_factor_on_a_pure_sponge_decay():
    """With g = 0, p = v = 0 and no PML, continuity reduces to
    rho' = -sponge(z) rho pointwise, so one STEP multiplies rho by the LDDRK
    polynomial R(-sponge dt) (step.f90:20-63); the zeroed planes k = 0 and
    nz-1 keep rho."""
    cfg = quiet_cfg(use_pml=False, npml=(0, 0, 0, 0, 0, 0), pulse=False, pulse_dir=0)
    bg = background(cfg, zbot=-1.0, ztop=0.3)
    bg["grav"] = np.zeros(cfg.nz)
    o = Oracle(cfg, bg)
    sp = o.array("spongexyz")
    assert sp.max() > 0
    a = np.zeros((5,) + sp.shape)
    rng = np.random.default_rng(2)
    a[0] = rng.random(sp.shape)
    o.put_state(a)
    o.step()
    zt = -sp * o.scalar("deltat")
    R = 1 + zt + zt ** 2 / 2 + 0.1665579725791151184 * zt ** 3 + 0.03950410250586729987 * zt ** 4 \
        + 0.007810706393337838236 * zt ** 5
    out = o.get_state()
    np.testing.assert_allclose(out[0][1:-1], (R * a[0])[1:-1], rtol=1e-13)
    np.testing.assert_array_equal(out[0][[0, -1]], a[0][[0, -1]])
    assert np.all(out[1:] == 0.0)


def test_filter_cadence():
    cfg = quiet_cfg()
    o = Oracle(cfg, background(cfg))
    rng = np.random.default_rng(3)
    a = rng.standard_normal((5, cfg.nz, cfg.ny, cfg.nx))
    o.put_state(a)
    o.filter_vars(cfg.zfreq * cfg.xyfreq + 1)      # neither cadence divides it
    np.testing.assert_array_equal(o.get_state(), a)
    o.filter_vars(cfg.zfreq)                          # z only (zfreq=3, xyfreq=5)
    b = o.get_state()
    assert not np.array_equal(b, a)
    # z-only filtering leaves each z column's DC content: rows 0 and nz-1 fixed
    np.testing.assert_array_equal(b[:, [0, -1]], a[:, [0, -1]])


# ----------------------------------------------------------------- set-up
def test_nondimensionalisation_and_metric():
    cfg = quiet_cfg()
    bg = background(cfg)
    o = Oracle(cfg, bg)
    assert o.scalar("dimc") == bg["cs0"][0, 0, 0]
    assert o.scalar("dimrho") == bg["rho0"][0, 0, 0]
    assert o.scalar("deltat") == cfg.timestep * o.scalar("dimc") / 69598946770.0
    assert o.scalar("stretchx") == 69598946770.0 / cfg.xlength * cfg.nx
    np.testing.assert_allclose(o.array("rho0")[:, 0, 0], bg["rho0"][:, 0, 0] / bg["rho0"][0, 0, 0])
    z = o.array("zaxis")
    np.testing.assert_allclose(z, 1 + bg["zaxis"] * 1e8 / 69598946770.0, rtol=1e-15)


def test_stabilise_enforces_convective_stability():
    """gradp0_z >= -rho0 g (and >= c^2 grad rho0 away from the surface band)."""
    cfg = quiet_cfg(use_pml=False, npml=(0, 0, 0, 0, 0, 0))
    o = Oracle(cfg, background(cfg))
    gp = o.array("gradp0_z")[:, 0, 0]
    gr = o.array("gradrho0_z")[:, 0, 0]
    c2 = o.array("c2")[:, 0, 0]
    rho0 = o.array("rho0")[:, 0, 0]
    g = o.array("g")
    z = o.array("zaxis")
    band = (z > np.float32(0.9998)) & (z < 1.0)
    rest = (z <= np.float32(0.9998)) | (z >= np.float32(1.0001))
    assert np.all(gp[rest] == np.maximum(c2 * gr, -rho0 * g)[rest])
    assert np.all(gp[band] >= (-rho0 * g)[band])


def test_z_pml_profiles():
    cfg = quiet_cfg(pmlc=((9.2103, 2.0, 3.0, 0.1),) * 6)
    o = Oracle(cfg, background(cfg))
    n = cfg.npml[4]
    k = o.array("pml_k4").reshape(n, cfg.ny, cfg.nx)[:, 0, 0]
    a = o.array("pml_a4").reshape(n, cfg.ny, cfg.nx)[:, 0, 0]
    b = o.array("pml_b4").reshape(n, cfg.ny, cfg.nx)[:, 0, 0]
    assert k[0] == pytest.approx(3.0) and k[-1] == 1.0     # kappa_max at the outer face
    assert a[-1] == 0.0 and a[0] < 0                        # damping vanishes at the inner edge
    assert b[0] < 0 and np.all(b < 0)
    g = o.array("g")
    assert g[0] == 0.0 and g[-1] == 0.0                     # alpha = 0 on the outer planes
    kt = o.array("pml_k5").reshape(cfg.npml[5], cfg.ny, cfg.nx)[:, 0, 0]
    assert kt[0] == 1.0 and kt[-1] == pytest.approx(3.0)


def test_alfven_limiter_caps_the_speed():
    cfg = mhd_cfg()
    o = Oracle(cfg, background(cfg))
    red = o.array("reduction")
    b2 = o.array("box") ** 2 + o.array("boy") ** 2 + o.array("boz") ** 2
    dimb, dimrho = o.scalar("dimb"), o.scalar("dimrho")
    va = np.sqrt(dimb ** 2 * red * b2 / (4 * math.pi * dimrho * o.array("rho0")))
    assert va.max() <= cfg.maxva * (1 + 1e-9)
    assert va.max() >= cfg.maxva * 0.999          # the search stops one decimal step short
    assert np.all((red > 0) & (red <= 1))


def test_source_known_answer():
    """At the source centre and t = t1 + t0/4 the carrier is sin(4.5 pi) = 1."""
    cfg = quiet_cfg(nx=32, ny=32, pulse_px=25e8 * 8 / 32, pulse_py=25e8 * 16 / 32,
                    pulse_pz=0.0, pulse_t0=40.0, pulse_t1=20.0, pulse_st=8.0, timestep=1.0)
    bg = background(cfg)
    o = Oracle(cfg, bg)
    t = 30                                  # t1 + t0/4
    o.forcing(t)
    f = o.array("forcing")
    z = o.array("zaxis")
    kc = int(np.argmin(np.abs(z - 1.0)))
    zz = (z[kc] - 1.0) / (cfg.pulse_sz / 69598946770.0)
    expect = (cfg.pulse_amp / 0.56458272 / o.scalar("dimc")
              * math.exp(-(10.0 ** 2) / (2 * 8.0 ** 2)) * math.exp(-zz ** 2 / 2))
    assert f[kc, 16, 8] == pytest.approx(expect, rel=1e-12)
    assert np.unravel_index(np.argmax(np.abs(f[kc])), f[kc].shape) == (16, 8)


def test_diag_known_answer():
    cfg = quiet_cfg()
    o = Oracle(cfg, background(cfg))
    a = np.zeros((5, cfg.nz, cfg.ny, cfg.nx))
    a[3] = 0.01                               # uniform vz
    o.put_state(a)
    d = o.diag()
    rho0 = o.array("rho0")
    m = rho0 * 0.01 ** 2 / 2
    expect = math.sqrt(np.sum(m * m) / a[0].size) * o.scalar("dimc") ** 2 * o.scalar("dimrho")
    assert d[0] == pytest.approx(expect, rel=1e-12)
