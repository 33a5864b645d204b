"""Parity on the benchmark workloads themselves, at their real sizes.

* quiet_acoustic (64 x 64 x 128, FP64): all 501 driver steps (maxtime 500,
  driver.f90:123-202) against the oracle, every field and CPML face within
  relative L2 1e-11, and the norm_KE history step by step.
* sunspot_mhd (256^3, FP64, 6-face CPML): three driver steps that include a
  filter on both cadences (steps 58..60: xy every 60, z every 20), from a
  seeded smooth state with seeded CPML bases, against the oracle (1e-11).
* sunspot_mhd at 64^3 with kappa = 1: 800 steps from rest through the onset
  of the slow growth the linearised run shows, compared with the oracle
  every 100 steps: the growth is the scheme's, reproduced to 1e-11.
* quiet_precision (512 x 512 x 300): FP32 + compensated update against
  FP64 after 500 steps, relative L2 <= 1e-5 per field.
* batch_population (the 64 seeded runs): batched equals alone, bit for bit,
  on a sample of the runs; the three smallest runs that stay within the
  failure checks, batched in FP32, stay within 1e-5 of their FP64 twins
  after 2000 steps.
* a Poisson request stream with the served form's grids (64^3 to
  384 x 384 x 256) at reduced step counts: each result equals the request
  served alone.
"""
import copy
import functools
import time

import numpy as np
import pytest

from oracle.oracle import Oracle
from tests.cases import field_errors, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-11


def _say(*a):
    print(*a, flush=True)


def smooth_state(cfg, st, seed, amp=1e-4):
    """Seeded smooth state scaled to the local background (rho by rho0, p
    by p0, v and b by the sound speed, so rho+rho0 and p+p0 stay positive
    in every layer): per field a sum of separable products of random
    low-order Fourier profiles (cheap at 256^3), plus seeded CPML bases."""
    rng = np.random.default_rng(seed)
    x = np.arange(cfg.nx) / cfg.nx
    y = np.arange(cfg.ny) / cfg.ny
    z = np.arange(cfg.nz) / cfg.nz
    a = np.zeros((cfg.nv, cfg.nz, cfg.ny, cfg.nx))

    def prof(t):
        m = rng.integers(1, 5)
        return np.cos(2 * np.pi * m * t + rng.uniform(0, 2 * np.pi))

    for q in range(cfg.nv):
        for _ in range(3):
            a[q] += rng.standard_normal() * (prof(z)[:, None, None] * prof(y)[None, :, None]
                                             * prof(x)[None, None, :])
    bg = st.bg
    shape = (cfg.nz, cfg.ny, cfg.nx)
    full = lambda v: np.broadcast_to(v if np.ndim(v) == 3 else np.reshape(v, (-1, 1, 1)), shape)
    c = np.sqrt(full(bg["c2"]))
    scale = [full(bg["rho0"]), c, c, c, full(bg["p0"])] + [c * np.sqrt(full(bg["rho0"]))] * 3
    for q in range(cfg.nv):
        a[q] *= amp * scale[q]
    psi = {f: amp * 0.1 * rng.standard_normal((cfg.npsi,) + cfg.face_shape(f))
           for f in range(6) if cfg.face_active(f)}
    return a, psi


@functools.lru_cache(maxsize=1)
def _population():
    from sparc2_amd.workloads import batch_population
    t0 = time.perf_counter()
    pop = batch_population(64)
    _say(f"population of 64 runs generated in {time.perf_counter() - t0:.1f} s")
    return pop


def _compare(s, o, label, with_psi=True):
    got, gpsi = s.get_state(with_psi=True)
    ref = o.get_state()
    errs = field_errors(got, ref)
    _say(f"{label}: " + " ".join(f"{e:.2e}" for e in errs))
    assert max(errs) < TOL
    if with_psi:
        for f, v in gpsi.items():
            e = rel_l2(v.ravel(), o.get_psi(f))
            assert e < TOL, (label, f, e)
    return errs


def test_quiet_acoustic_full_run():
    from sparc2_amd import workloads
    from sparc2_amd.solver import Solver
    cfg, bg = workloads.build("quiet_acoustic")
    assert (cfg.nx, cfg.ny, cfg.nz, cfg.maxtime) == (64, 64, 128, 500)
    o = Oracle(cfg, bg)
    t0 = time.perf_counter()
    hist_o = o.run(0, cfg.maxtime)
    _say(f"oracle: {cfg.maxtime + 1} steps in {time.perf_counter() - t0:.1f} s")
    s = Solver(cfg, bg)
    hist = s.run(0, cfg.maxtime)
    assert hist.shape == hist_o.shape
    np.testing.assert_allclose(hist[:, 0], hist_o[:, 0], rtol=1e-9)
    np.testing.assert_allclose(hist[:, 1:], hist_o[:, 1:], rtol=1e-12)
    assert hist_o[-1, 0] > 0
    _compare(s, o, "quiet_acoustic 501 steps")
    s.close()


def test_sunspot_mhd_full_grid_filter_steps():
    from sparc2_amd import workloads
    from sparc2_amd.solver import Solver
    cfg, bg = workloads.build("sunspot_mhd")
    assert (cfg.nx, cfg.ny, cfg.nz) == (256, 256, 256)
    assert 60 % cfg.xyfreq == 0 and 60 % cfg.zfreq == 0
    s = Solver(cfg, bg)
    a, psi = smooth_state(cfg, s.st, seed=31)
    o = Oracle(cfg, bg)
    o.put_state(a)
    for f, v in psi.items():
        o.put_psi(f, v)
    t0 = time.perf_counter()
    hist = o.run(58, 60)
    _say(f"oracle: 3 steps of 256^3 MHD in {time.perf_counter() - t0:.1f} s; diag {hist[-1]}")
    s.put_state(a, psi)
    s.advance(58, 60)
    _compare(s, o, "sunspot_mhd 256^3 steps 58..60")
    # the last step's diagnostics (before its filter, driver.f90:161-183)
    d = s.diag()
    assert abs(d[0] - hist[-1, 0]) <= 1e-10 * abs(hist[-1, 0])
    np.testing.assert_allclose(d[1:], hist[-1, 1:], rtol=1e-12)
    s.close()


def test_sunspot_mhd_kappa1_through_growth():
    from sparc2_amd import workloads
    from sparc2_amd.solver import Solver
    cfg, bg = workloads.build("sunspot_mhd", nx=64, ny=64, nz=64)
    assert cfg.pmlc[0][2] == 1.0
    o = Oracle(cfg, bg)
    s = Solver(cfg, bg)
    norms = []
    for n0 in range(0, 800, 100):
        o.run(n0, n0 + 99)
        s.advance(n0, n0 + 99)
        _compare(s, o, f"64^3 MHD kappa=1 after step {n0 + 99}")
        norms.append(o.diag()[0])
    _say("norm_KE every 100 steps:", " ".join(f"{v:.3e}" for v in norms))
    s.close()


def test_quiet_precision_fp32_against_fp64():
    from sparc2_amd import workloads
    from sparc2_amd.solver import Solver
    cfg64, bg = workloads.build("quiet_precision", precision="fp64")
    assert (cfg64.nx, cfg64.ny, cfg64.nz) == (512, 512, 300)
    s64 = Solver(cfg64, bg)
    s32 = Solver(cfg64.replace(precision="fp32"), bg, setup=s64.st)
    s64.advance(0, 499)
    s32.advance(0, 499)
    a64 = s64.get_state()
    a32 = s32.get_state()
    errs = field_errors(a32, a64)
    _say("quiet_precision FP32 vs FP64 after 500 steps: " + " ".join(f"{e:.2e}" for e in errs))
    assert max(errs) < 1e-5
    s64.close()
    s32.close()


def test_batch_population_equals_alone():
    from sparc2_amd.batch import Batch
    from sparc2_amd.solver import Solver
    pop = _population()
    t0 = time.perf_counter()
    sol = [Solver(c, b) for c, b in pop]
    _say(f"64 runs set up in {time.perf_counter() - t0:.1f} s; "
         f"{sum(c.mhd for c, _ in pop)} MHD, {sum(c.npoints for c, _ in pop) / 1e6:.1f} M points")
    n = 40
    b = Batch(sol, use_graph=True)
    b.advance(n)
    sample = [0, 17, 38, 63]
    got = {q: sol[q].get_state(with_psi=True) for q in sample}
    b.close()
    for s in sol:
        s.close()
    for q in sample:
        c, bgq = pop[q]
        s = Solver(c, bgq)
        s.advance(0, n - 1)
        a, psi = s.get_state(with_psi=True)
        np.testing.assert_array_equal(got[q][0], a)
        for f in psi:
            np.testing.assert_array_equal(got[q][1][f], psi[f])
        _say(f"run {q} ({c.nx}x{c.ny}x{c.nz}, {'MHD' if c.mhd else 'quiet'}): batch == alone")
        s.close()


def test_batch_population_run_fp32_after_2000_steps():
    """The three smallest population runs that stay within the reference's
    failure checks for 2000 FP64 steps, batched in FP32, against the same
    runs alone in FP64 (which equal the oracle to 1e-11) after 2000 steps:
    relative L2 <= 1e-5 per field.  Several small flux-tube runs grow until
    p + p0 turns negative after 1400-1850 s of simulated time whatever the
    time step (the linearised equations about these generated backgrounds
    are unstable; the oracle grows alike), so runs are taken in size order
    and those whose FP64 run trips the checks are passed over."""
    from sparc2_amd._lib import GUARD_RECORD
    from sparc2_amd.batch import Batch
    from sparc2_amd.solver import Solver
    pop = _population()
    pick, ref = [], {}
    for q in sorted(range(64), key=lambda q: pop[q][0].npoints):
        c, bgq = pop[q]
        s64 = Solver(c.replace(precision="fp64"), bgq)
        s64.set_guard(GUARD_RECORD)          # note the first failure, keep stepping
        s64.advance(0, 1999, check_guard=False)
        failed, t, _, _ = s64.guard()
        _say(f"run {q} ({c.nx}x{c.ny}x{c.nz}) FP64 2000 steps: " +
             (f"fails the checks at step {t}, passed over" if failed else "within the checks"))
        if failed:
            s64.close()
            continue
        pick.append(q)
        ref[q] = s64
        if len(pick) == 3:
            break
    assert len(pick) == 3
    sol = [Solver(pop[q][0], pop[q][1], setup=ref[q].st) for q in pick]
    b = Batch(sol, use_graph=True)
    assert not b.advance(2000)
    for q, s32 in zip(pick, sol):
        c = pop[q][0]
        s64 = ref[q]
        errs = field_errors(s32.get_state(), s64.get_state())
        _say(f"run {q} ({c.nx}x{c.ny}x{c.nz}, {'MHD' if c.mhd else 'quiet'}) FP32 batched vs FP64 "
             f"after 2000 steps: " + " ".join(f"{e:.1e}" for e in errs))
        assert max(errs) < 1e-5
        s64.close()
    b.close()
    for s in sol:
        s.close()


def test_served_stream_full_grids_mix_equals_alone():
    from sparc2_amd.serve import Server, poisson_stream
    reqs = poisson_stream(5, seed=11, steps=(20, 120))
    for r in reqs:
        r.slice_every_s = 4.0          # a few slices within the short runs
    sizes = sorted({(r.cfg.nx, r.cfg.ny, r.cfg.nz) for r in reqs})
    _say("stream grids:", sizes, "steps:", [r.nsteps for r in reqs])
    mixed = Server(max_chunk=40, keep_final=True).run(copy.deepcopy(reqs), verbose=True)
    for r in reqs:
        alone = Server(max_chunk=40, keep_final=True).run([copy.deepcopy(r)])
        a, b = mixed["results"][r.rid], alone["results"][r.rid]
        assert a.steps == b.steps == r.nsteps
        assert len(a.slices) == len(b.slices)
        for (ta, sa), (tb, sb) in zip(a.slices, b.slices):
            assert ta == tb
            np.testing.assert_array_equal(sa, sb)
        np.testing.assert_array_equal(a.final, b.final)
        _say(f"request {r.rid} {r.cfg.nx}x{r.cfg.ny}x{r.cfg.nz}: mix == alone")
    _say({k: v for k, v in mixed.items() if k != "results"})
