"""SSP33!'s history output, check_conservation and checkpoints on the GPU, through the C-ABI.

* ``SSP33`` with ``output_interval`` (p2de_run_segment) against the oracle stepping the same
  loop (src/timestepping/SSPRK33.jl:28-56): the same history entries (a step i with
  mod(i, output_interval) == 0 or |t - T| < 1e-10), every dt to 1e-13, the entries' times,
  states, L (Zhang-Shu element parameters, K x 3) and theta (element means of theta_local,
  K x 3), and check_conservation of the device state against the host sum of the oracle's.
  States, L and theta meet the default bar (1e-12; 1e-6 for the min-entropy bound, see
  tests/test_gpu_limiters.py) or 100 times the oracle's own sensitivity, whichever is
  larger: the oracle run a second time from the initial state perturbed by 1e-15 (relative,
  seeded) departs from the first by that much at each entry.  On the coarse 5x5 smoke mesh
  the unlimited Gauss scheme (Zhang-Shu keeps l = 1 there) amplifies such a perturbation
  about tenfold per step (1.5e-8 after 11 steps); the GPU's per-step rounding grows alike;
* a checkpoint taken after n1 steps and loaded into a fresh context continues bit for bit
  as the run that was not stopped (also for the min-entropy bounds, whose t0 minimum the
  checkpoint carries, and for the mixed-precision form); a corrupted, truncated or foreign
  checkpoint is rejected and leaves the context as it was;
* the Shu-Osher example (examples/1D/shu-osher.jl: Gauss, N = 3, K = 64, T = 1.8, inflow and
  outflow) for 36000 steps: the GPU's density against the oracle's and against the WENO5
  solution of data/weno5_shuosher.mat (tests/golden/weno5_shuosher.npz).
"""
import dataclasses

import numpy as np
import pytest

from conftest import PKG, normwise_rel
from oracle.oracle import Oracle
from test_postprocess import weno5

pytestmark = pytest.mark.gpu

P = PKG.postprocess


def build(name, output_interval, upper_bound_mode=0, **kw):
    param, ic, bc = PKG.configs.CONFIGS[name](**kw)
    param = dataclasses.replace(param, postprocessing_param=PKG.PostprocessingParameter(output_interval=output_interval))
    solver, state, sp = PKG.initialize_DG(param, ic, bc)
    fp = PKG.flat_problem(solver, sp, upper_bound_mode=upper_bound_mode)
    return solver, state, sp, fp


def oracle_history(solver, fp, U, interval):
    """SSPRK33.jl:24-56 with the oracle's step: dthist and the history entries."""
    orc = Oracle(fp)
    T = fp.T
    t, i = fp.t0, 1
    h = PKG.DataHistory()
    while t < T:
        t, dt = orc.step(U, t)
        i += 1
        h.dthist.append(dt)
        if i % interval == 0 or abs(t - T) < 1e-10:
            h.thist.append(t)
            h.Uhist.append(U.copy())
            L = orc.field("L_zhangshu").reshape(3, fp.K).T if fp.rhs_limiter == 1 else np.zeros((fp.K, 3))
            if fp.entropyproj_limiter == 1:
                th = orc.field("theta_local").reshape(3, fp.K, fp.Nfp).mean(axis=2).T
            else:
                th = np.zeros((fp.K, 3))
            h.Lhist.append(L)
            h.thetahist.append(th)
            h.conservation.append(P.check_conservation(U, solver))
    return h


HISTORY_CASES = [("subcell", "lobatto"), ("zhangshu", "lobatto"), ("subcell", "gauss"), ("zhangshu", "gauss"),
                 ("minentropy", "gauss")]


@pytest.mark.parametrize("case", HISTORY_CASES, ids=["-".join(c) for c in HISTORY_CASES])
def test_ssp33_history_matches_the_oracle_loop(case):
    limiter, basis = case
    solver, state, sp, fp = build("smoke_vortex", 3, N=3, T=0.1, limiter=limiter, basis=basis)
    print(f"case smoke_vortex N=3 T=0.1 output_interval=3 limiter {limiter} basis {basis}", flush=True)
    U0 = state.preallocation.Uq.copy()
    ref = oracle_history(solver, fp, U0.copy(), 3)
    rng = np.random.default_rng(7)
    twin = oracle_history(solver, fp, U0 * (1 + 1e-15 * rng.uniform(-1, 1, U0.shape)), 3)
    hist = PKG.SSP33(state, solver, sp)
    base = 1e-6 if fp.bound != 0 else 1e-12
    print(f"  {len(hist.dthist)} steps, entries at t = {hist.thist}", flush=True)
    assert len(hist.dthist) == len(ref.dthist) and len(hist.thist) == len(ref.thist) >= 3
    np.testing.assert_allclose(hist.dthist, ref.dthist, rtol=1e-13, atol=0)
    np.testing.assert_allclose(hist.thist, ref.thist, rtol=1e-13, atol=0)
    assert abs(hist.thist[-1] - fp.T) < 1e-10
    for n in range(len(ref.thist)):
        e = normwise_rel(hist.Uhist[n], ref.Uhist[n], fp.Nc)
        eL = float(np.max(np.abs(hist.Lhist[n] - ref.Lhist[n])))
        eth = float(np.max(np.abs(hist.thetahist[n] - ref.thetahist[n])))
        ec = abs(hist.conservation[n] - ref.conservation[n]) / abs(ref.conservation[n])
        sens = normwise_rel(twin.Uhist[n], ref.Uhist[n], fp.Nc)
        tol = max(base, 100 * sens)
        print(f"  entry {n}: t {hist.thist[n]:.6f} state {e:.2e} (oracle sensitivity {sens:.1e}) L {eL:.2e} "
              f"theta {eth:.2e} conservation {hist.conservation[n]:.17g} ({ec:.1e})", flush=True)
        assert e <= tol
        assert eL <= tol and eth <= tol
        assert ec <= 1e-13
    if fp.rhs_limiter != 1:
        assert not np.any(hist.Lhist[-1])                # the subcell limiter never writes L
    if fp.entropyproj_limiter == 1:
        assert np.all((hist.thetahist[-1] > 0) & (hist.thetahist[-1] <= 1))
    else:
        assert not np.any(hist.thetahist[-1])
    assert np.array_equal(state.preallocation.Uq, hist.Uhist[-1])
    assert hist.positivity_violations == 0


def test_run_segment_stops_at_entries_and_limits():
    from p2de_jl_amd.solver import DeviceSolver
    solver, state, sp, fp = build("smoke_vortex", 4, N=2, T=0.1)
    dev = DeviceSolver(fp)
    dev.set_state(state.preallocation.Uq)
    dev.set_time(0.0)
    dts, i, out = dev.run_segment(4, 1)
    assert (len(dts), i, out) == (3, 4, True)                  # steps 1..3: i = 2, 3, 4
    dts, i, out = dev.run_segment(4, i, max_steps=2)
    assert (len(dts), i, out) == (2, 6, False)                 # stopped by max_steps
    dts, i, out = dev.run_segment(4, i, cap=1)
    assert (len(dts), i, out) == (1, 7, False)                 # stopped by the dt buffer
    dts, i, out = dev.run_segment(4, i)
    assert (len(dts), i, out) == (1, 8, True)
    t0, _ = dev.get_time()
    # the plain loop is unaffected by a finished segment
    st = dev.run(max_steps=3)
    assert st.steps == 3 and st.t > t0
    with pytest.raises(PKG.P2DEError):
        dev.run_segment(0, 1)


CKPT_CASES = [("subcell", "lobatto"), ("relaxed_minentropy", "gauss"), ("zhangshu", "gauss"),
              ("minentropy", "lobatto")]


@pytest.mark.parametrize("case", CKPT_CASES, ids=["-".join(c) for c in CKPT_CASES])
def test_checkpoint_resume_is_bitwise(case, tmp_path):
    from p2de_jl_amd.solver import DeviceSolver
    limiter, basis = case
    solver, state, sp, fp = build("smoke_vortex", 10**9, N=3, T=1.0, limiter=limiter, basis=basis)
    print(f"case smoke_vortex N=3 limiter {limiter} basis {basis}: checkpoint after 3 steps, 4 more", flush=True)
    a = DeviceSolver(fp)
    a.set_state(state.preallocation.Uq)
    a.set_time(0.0)
    assert a.run(max_steps=3).steps == 3
    P.save_checkpoint(a, tmp_path / "ck.bin", step_index=4)
    assert a.run(max_steps=4).steps == 4
    Ua, ta = a.get_state(), a.get_time()
    b = DeviceSolver(fp)
    assert P.load_checkpoint(b, tmp_path / "ck.bin") == 4
    assert b.run(max_steps=4).steps == 4
    Ub, tb = b.get_state(), b.get_time()
    print(f"  t {ta} vs {tb}, max |dU| {np.max(np.abs(Ua - Ub)):.3e}", flush=True)
    assert ta == tb
    assert np.array_equal(Ua, Ub)


def test_checkpoint_rejects_bad_blocks():
    from p2de_jl_amd.solver import DeviceSolver
    solver, state, sp, fp = build("smoke_vortex", 10**9, N=2, T=1.0)
    dev = DeviceSolver(fp)
    dev.set_state(state.preallocation.Uq)
    dev.set_time(0.0)
    dev.run(max_steps=1)
    blob = dev.checkpoint(2)
    U1, t1 = dev.get_state(), dev.get_time()
    bad = bytearray(blob)
    bad[-3] ^= 0x10                                             # a flipped bit in the state
    _, _, _, fp3 = build("smoke_vortex", 10**9, N=3, T=1.0)
    other = DeviceSolver(fp3)
    other.set_state(PKG.configs.build("smoke_vortex", N=3)[1].preallocation.Uq)
    for blk, what in [(bytes(bad), "checksum"), (blob[:-8], "truncated"), (blob[:40], "truncated"),
                      (b"X" + blob[1:], "magic"), (other.checkpoint(1), "another problem")]:
        with pytest.raises(PKG.P2DEError) as ei:
            dev.restore(blk)
        print(f"  {what}: {ei.value}", flush=True)
        assert ei.value.status == 1
    assert np.array_equal(dev.get_state(), U1) and dev.get_time() == t1
    assert dev.restore(blob) == 2


def test_mixed_precision_history_and_checkpoint():
    from p2de_jl_amd.solver import DeviceSolver, PRECISION_MIXED
    solver, state, sp, fp = build("vortex_2d", 2, K=8, T=1.0)
    dev = DeviceSolver(fp, precision=PRECISION_MIXED)
    dev.set_state(state.preallocation.Uq)
    dev.set_time(0.0)
    dts, i, out = dev.run_segment(2, 1)
    assert (len(dts), i, out) == (1, 2, True) and dts[0] > 0
    dts, i, out = dev.run_segment(2, i, max_steps=5)
    assert (len(dts), i, out) == (2, 4, True)
    blob = dev.checkpoint(i)
    dev.run(max_steps=3)
    Ua = dev.get_state()
    fresh = DeviceSolver(fp, precision=PRECISION_MIXED)
    assert fresh.restore(blob) == 4
    fresh.run(max_steps=3)
    assert np.array_equal(fresh.get_state(), Ua) and fresh.get_time() == dev.get_time()
    assert not np.any(dev.field(PKG.solver.FIELD_L)) and not np.any(dev.field(PKG.solver.FIELD_THETA_MEAN))
    tot = dev.check_conservation()
    assert tot == pytest.approx(P.check_conservation(Ua, solver), rel=1e-13)


@pytest.mark.parametrize("mode", [1, 0], ids=["intended-upper-bound", "reference-upper-bound"])
def test_shu_osher_against_oracle_and_weno5(mode):
    """36000 steps of examples/1D/shu-osher.jl on the GPU through SSP33 with output every 1000
    steps; the final density against the oracle's run and against WENO5."""
    solver, state, sp, fp = build("shu_osher", 1000, upper_bound_mode=mode)
    print(f"case shu_osher K=64 N=3 Gauss upper_bound_mode {mode}", flush=True)
    U_cpu = state.preallocation.Uq.copy()
    t = 0.0
    orc = Oracle(fp)
    while t < fp.T:
        t, _ = orc.step(U_cpu, t)
    from p2de_jl_amd.solver import DeviceSolver
    state.device = DeviceSolver(fp)
    hist = PKG.SSP33(state, solver, sp)
    U = state.preallocation.Uq
    xw, rw = weno5()
    x = solver.md.xq.T.ravel()
    rho = U[:, :, 0].ravel()
    ri = np.interp(x, xw, rw)
    l1 = np.sum(np.abs(rho - ri)) / np.sum(np.abs(ri))
    e = normwise_rel(U, U_cpu, fp.Nc)
    print(f"  {len(hist.dthist)} steps, {len(hist.thist)} entries, GPU vs oracle {e:.2e}, L1 vs WENO5 {l1:.4f}",
          flush=True)
    # i = 1000, 2000, ..., 36000 (step 35999) and the last step, where |t - T| < 1e-10
    assert len(hist.dthist) == 36000 and len(hist.thist) == 37
    assert abs(hist.thist[-1] - 1.8) < 1e-10
    assert e <= 1e-10        # the oracle itself departs by 3.4e-13 (1.4e-14) under a 1e-15 initial perturbation
    assert hist.positivity_violations == 0
    if mode == 1:
        assert l1 < 0.015
