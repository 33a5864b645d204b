"""The reference's time-loop schedule on the GPU path (driver.f90:123-202):
outputs after STEP and before FILTER_VARS, restart from a snapshot without
filtering its step again, and the per-step failure checks that stop a run
at the failing step and dump it to full_XXXX.h5 -- each compared with the
oracle driven the same way.  FP64 results within relative L2 1e-11 of the
oracle; FP32 within 1e-5."""
import os

import numpy as np
import pytest

from oracle.oracle import Oracle
from sparc2_amd import h5io
from tests.cases import background, field_errors, mhd_cfg, quiet_cfg, random_state, rel_l2

pytestmark = pytest.mark.gpu
H5 = pytest.mark.skipif(not os.path.exists(h5io.LIB), reason="libsparc_h5.so not built")

TOL = 1e-11


def _say(*a):
    print(*a, flush=True)


def _cfg(case, **kw):
    return {"quiet": lambda: quiet_cfg(**kw), "mhd": lambda: mhd_cfg(**kw)}[case]()


def _oracle_steps(o, n0, n1, filter_last=True):
    """driver.f90:123-200 on the oracle, the filter of n1 optional."""
    for t in range(n0, n1 + 1):
        o.forcing(t)
        o.step()
        if t != 0 and (t < n1 or filter_last):
            o.filter_vars(t)


def _file_state(path, cfg, st):
    shape = (cfg.nz, cfg.ny, cfg.nx)
    a = np.empty((cfg.nv,) + shape)
    for name, f, _, from_file in h5io._scales(st, cfg.nv):
        a[f] = from_file(h5io.read(path, name, shape))
    return a


@H5
@pytest.mark.parametrize("case", ["quiet", "mhd"])
def test_outputs_are_pre_filter_and_restart_skips_the_snapshot_filter(tmp_path, case):
    """Snapshots at a step on both filter cadences hold the state before its
    filter; a run restarted from one continues at time0*fulloutputsteps + 1
    without filtering the snapshot's step again (driver.f90:61-69,198-200),
    and equals the oracle doing the same."""
    from sparc2_amd.driver import restart, run_driver
    from sparc2_amd.solver import Solver
    cfg = _cfg(case)
    full = cfg.xyfreq * cfg.zfreq                   # a step on both cadences
    assert full > 1 and full % cfg.xyfreq == 0 and full % cfg.zfreq == 0
    bg = background(cfg)
    a, psi = random_state(cfg, seed=41)
    s = Solver(cfg, bg)
    s.put_state(a, psi)
    out = run_driver(s, 0, 2 * full + 2, out_dir=str(tmp_path), simname="run_",
                     fulloutputsteps=full, chunkoutputsteps=cfg.zfreq)
    kinds = [(k, t) for k, t, _ in out]
    assert ("full", 0) in kinds and ("full", full) in kinds and ("full", 2 * full) in kinds
    assert ("chunk", cfg.zfreq) in kinds and ("chunk", 0) not in kinds
    o = Oracle(cfg, bg)
    o.put_state(a)
    for f, v in psi.items():
        o.put_psi(f, v)
    _oracle_steps(o, 0, full, filter_last=False)
    snap = os.path.join(str(tmp_path), h5io.file_name("run_", "full", 1))
    errs = field_errors(_file_state(snap, cfg, s.st), o.get_state())
    _say(f"{case} snapshot at {full} vs oracle before the filter: " + " ".join(f"{e:.1e}" for e in errs))
    assert max(errs) < 1e-13
    # restart from it: the snapshot's filter is not applied (time 'full' is
    # done), the run goes on at full + 1
    r = Solver(cfg, bg, setup=s.st)
    t1 = restart(r, snap, full)
    assert t1 == full + 1
    run_driver(r, t1, full + 9)
    _oracle_steps(o, full + 1, full + 9)
    got, gpsi = r.get_state(with_psi=True)
    errs = field_errors(got, o.get_state())
    _say(f"{case} restarted at {t1} vs oracle after {full + 9}: " + " ".join(f"{e:.1e}" for e in errs))
    assert max(errs) < TOL
    for f, v in gpsi.items():
        assert rel_l2(v.ravel(), o.get_psi(f)) < TOL
    s.close()
    r.close()


@H5
def test_fp32_restart_keeps_the_compensation(tmp_path):
    """An FP32 run written to full_NNNN.h5 (the compensated value a - c in
    FP64) and read back continues as the uninterrupted run: put_state
    splits the value into (fl32(x), fl32(fl32(x) - x)), which represents x
    to FP64 accuracy.  After filters the uninterrupted run's pair can hold
    the same value split differently, so the stage values (a + beta K,
    rounded to FP32) may differ by an FP32 ulp: the two continuations agree
    to FP32 rounding (measured 4.7e-8; bar 1e-6), far inside the FP32 form's
    1e-5."""
    from sparc2_amd.solver import Solver
    cfg = quiet_cfg(precision="fp32")
    bg = background(cfg)
    a, psi = random_state(cfg, seed=42)
    s = Solver(cfg, bg)
    s.put_state(a, psi)
    t = 7                                   # on neither filter cadence (3, 5)
    assert t % cfg.zfreq and t % cfg.xyfreq
    s.advance(0, t)
    p = str(tmp_path / h5io.file_name("run_", "full", 1))
    h5io.write_full(p, s, time=t)
    s.advance(t + 1, t + 12)
    ref = s.get_state()
    r = Solver(cfg, bg, setup=s.st)
    h5io.read_full(p, r)
    r.advance(t + 1, t + 12)
    errs = field_errors(r.get_state(), ref)
    _say("fp32 restart vs uninterrupted: " + " ".join(f"{e:.1e}" for e in errs))
    assert max(errs) < 1e-6
    s.close()
    r.close()


def _failing_state(cfg, seed):
    """A state whose density perturbation makes rho + rho0 negative at one
    point (the reference's check trips at the first step)."""
    from sparc2_amd.initialise import prepare
    a, psi = random_state(cfg, seed=seed)
    st = prepare(cfg, background(cfg))
    rho0 = st.bg["rho0"]
    k, j, i = cfg.nz // 2, cfg.ny // 3, cfg.nx // 4
    a[0, k, j, i] = -3.0 * (rho0[k] if np.ndim(rho0) == 1 else rho0[k, j, i])
    return a, psi


@H5
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_guard_stops_the_run_at_the_failing_step(tmp_path, precision):
    """min(rho+rho0) <= 0 at time index 5 (a filter step): advance raises
    DivergenceError naming step 5, the state stays the one after STEP 5 and
    before its filter -- later steps and filters leave it alone -- and the
    dump full_XXXX.h5 holds it (driver.f90:175-196, all_modules.f90:1090-1100)."""
    from sparc2_amd.solver import DivergenceError, Solver
    cfg = quiet_cfg(precision=precision)
    assert 5 % cfg.xyfreq == 0
    bg = background(cfg)
    a, psi = _failing_state(cfg, 43)
    s = Solver(cfg, bg)
    s.put_state(a, psi)
    with pytest.raises(DivergenceError) as ei:
        s.advance(5, 12)
    assert ei.value.time == 5 and ei.value.diag[1] <= 0.0
    got = s.get_state()
    o = Oracle(cfg, bg)
    o.put_state(a)
    for f, v in psi.items():
        o.put_psi(f, v)
    _oracle_steps(o, 5, 5, filter_last=False)
    tol = TOL if precision == "fp64" else 1e-5
    errs = field_errors(got, o.get_state())
    _say(f"{precision} halted state vs oracle after STEP 5: " + " ".join(f"{e:.1e}" for e in errs))
    assert max(errs) < tol
    np.testing.assert_allclose(ei.value.diag[1], o.diag()[1], rtol=1e-6 if precision == "fp32" else 1e-12)
    # further calls leave the halted state alone
    s.advance(13, 20, check_guard=False)
    s.filter_vars(15)
    np.testing.assert_array_equal(s.get_state(), got)
    with pytest.raises(DivergenceError) as e2:
        s.check_guard(str(tmp_path), "run_")
    dump = e2.value.dump
    assert dump and os.path.basename(dump) == "run_full_XXXX.h5"
    _, hours = h5io.attrs(dump)
    assert hours == pytest.approx(5 * cfg.timestep / 3600.0, rel=1e-14)
    errs = field_errors(_file_state(dump, cfg, s.st), got)
    assert max(errs) < 1e-12
    # a fresh state clears the guard; record mode notes a failure and runs on
    s.put_state(a, psi)
    s.set_guard(1)
    s.advance(5, 9, check_guard=False)
    failed, t, d, n = s.guard()
    assert failed and t == 5 and n == 8 + 8 + 5
    s.close()


def test_guard_record_mode_keeps_stepping():
    from sparc2_amd.solver import Solver
    cfg = quiet_cfg()
    bg = background(cfg)
    a, psi = _failing_state(cfg, 44)
    s = Solver(cfg, bg)
    s.set_guard(1)
    s.put_state(a, psi)
    s.advance(0, 9, check_guard=False)
    failed, t, d, n = s.guard()
    assert failed and t == 0
    o = Oracle(cfg, bg)
    o.put_state(a)
    for f, v in psi.items():
        o.put_psi(f, v)
    _oracle_steps(o, 0, 9)
    errs = field_errors(s.get_state(), o.get_state())
    assert max(errs) < TOL
    s.close()


def test_batch_failing_run_stops_alone_and_outputs_are_pre_filter():
    """A batch of three runs, one of which fails its checks at its first
    step: advance reports only that run (stopped at that step, equal to the
    oracle), the others equal their runs alone.  With defer_filter the runs
    hold the state after STEP and before the filter of the last step
    (driver.f90:198-200) until filter_deferred()."""
    from sparc2_amd.batch import Batch
    from sparc2_amd.solver import Solver
    cfgs = [quiet_cfg(), mhd_cfg(), quiet_cfg(nx=48, ny=32, nz=40, pulse_px=20e8)]
    bgs = [background(c) for c in cfgs]
    states = [random_state(cfgs[0], 45), _failing_state(cfgs[1], 46), random_state(cfgs[2], 47)]
    sol = []
    for c, b, (a, psi) in zip(cfgs, bgs, states):
        s = Solver(c, b)
        s.put_state(a, psi)
        sol.append(s)
    B = Batch(sol, use_graph=True)
    nsteps = cfgs[0].xyfreq * cfgs[0].zfreq + 1           # the last step filters run 0 on both cadences
    failed = B.advance(nsteps, defer_filter=True)
    assert list(failed) == [1] and failed[1].time == 0
    pre = [s.get_state() for s in sol]
    B.filter_deferred()
    post = [s.get_state() for s in sol]
    for q in (0, 2):
        o = Oracle(cfgs[q], bgs[q])
        o.put_state(states[q][0])
        for f, v in states[q][1].items():
            o.put_psi(f, v)
        _oracle_steps(o, 0, nsteps - 1, filter_last=False)
        assert max(field_errors(pre[q], o.get_state())) < TOL
        o.filter_vars(nsteps - 1)
        assert max(field_errors(post[q], o.get_state())) < TOL
    o = Oracle(cfgs[1], bgs[1])
    o.put_state(states[1][0])
    for f, v in states[1][1].items():
        o.put_psi(f, v)
    _oracle_steps(o, 0, 0, filter_last=False)
    assert max(field_errors(post[1], o.get_state())) < TOL
    B.close()
    for s in sol:
        s.close()
