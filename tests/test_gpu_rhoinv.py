"""The FP64 MHD xy kernel forms rhoinv as 1/rho0 (IEEE division, correctly
rounded) instead of reading the array when the host's rhoinv holds exactly
that at every point, as the reference's set-up makes it
(initialize.f90:1012).  The state after N steps equals, bitwise, the state
computed reading the array (SPARC_RHOINV=load); a host array that differs
from 1/rho0 anywhere is read, not replaced."""
import os

import numpy as np
import pytest

from tests.cases import background, mhd_cfg, random_state

pytestmark = pytest.mark.gpu


def _run(cfg, bg, a, psi, n, load):
    from sparc2_amd.solver import Solver
    old = os.environ.get("SPARC_RHOINV")
    if load:
        os.environ["SPARC_RHOINV"] = "load"
    elif old is not None:
        del os.environ["SPARC_RHOINV"]
    try:
        s = Solver(cfg, bg)
    finally:
        if old is None:
            os.environ.pop("SPARC_RHOINV", None)
        else:
            os.environ["SPARC_RHOINV"] = old
    s.put_state(a, psi)
    s.advance(0, n)
    out = s.get_state(with_psi=True)
    s.close()
    return out


@pytest.mark.parametrize("periodic", [False, True])
def test_rhoinv_formed_in_kernel_equals_loaded(periodic):
    cfg = mhd_cfg(nx=48, ny=40, nz=40, periodic=periodic)
    bg = background(cfg)
    a, psi = random_state(cfg, seed=21)
    print(f"rhoinv periodic={periodic}", flush=True)
    (s1, p1) = _run(cfg, bg, a, psi, 10, load=False)
    (s2, p2) = _run(cfg, bg, a, psi, 10, load=True)
    assert np.abs(s1[1]).max() > 0
    np.testing.assert_array_equal(s1, s2)
    for f in p2:
        np.testing.assert_array_equal(p1[f], p2[f])


def test_rhoinv_other_than_inverse_is_read():
    """A host rhoinv that is not 1/rho0 at one point: the kernel reads the
    array (the state equals the SPARC_RHOINV=load state bitwise and differs
    from the unperturbed run)."""
    import copy

    from sparc2_amd.solver import Solver
    cfg = mhd_cfg(nx=48, ny=40, nz=40)
    bg = background(cfg)
    a, psi = random_state(cfg, seed=22)
    s0 = Solver(cfg, bg)
    st = copy.copy(s0.st)
    s0.close()
    st.bg = dict(st.bg)
    ri = np.array(st.bg["rhoinv"], dtype=np.float64, copy=True)
    ri[20, 17, 23] *= 1.001
    st.bg["rhoinv"] = ri

    def run(setup, load):
        old = os.environ.pop("SPARC_RHOINV", None)
        if load:
            os.environ["SPARC_RHOINV"] = "load"
        try:
            s = Solver(cfg, bg, setup=setup)
        finally:
            os.environ.pop("SPARC_RHOINV", None)
            if old is not None:
                os.environ["SPARC_RHOINV"] = old
        s.put_state(a, psi)
        s.advance(0, 6)
        out = s.get_state()
        s.close()
        return out

    print("rhoinv perturbed", flush=True)
    got, ref = run(st, False), run(st, True)
    np.testing.assert_array_equal(got, ref)
    plain = _run(cfg, bg, a, psi, 6, load=False)[0]
    assert np.abs(plain - got).max() > 0
