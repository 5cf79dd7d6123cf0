"""Edge cases of the mesh and the time loop on the GPU, against the oracle:

* a single periodic element (every face its own opposite face's neighbour), a one-element-wide
  strip (Kx = 1), a non-square element grid with a near-vacuum dip (min rho 1e-6, anisotropic
  metric), the Gauss basis on a 3x2 grid with a dip, and 1D with one and two elements
  (transmissive ends; the reference's 1D symmetrisation wraps around, here onto the element
  itself): rhs parity to 1e-11 (worst node and 2-norm, per component) for three stages on the
  initial and a perturbed state, and three SSP-RK3 steps to 100 times the oracle's own
  sensitivity (1e-12 at least).  The 2-norm bar is the worst-node one here: the logarithmic mean
  of two values a ratio 1 + f apart (compressible_Navier_Stokes.jl:303-321) takes
  (b - a)/(log b - log a) for f >= 1e-4, where one ulp of log a (the GPU's and glibc's
  faithful logs may differ by one) moves the mean by ulp(log a)/f, up to 4e-12 relative at
  f = 1e-4 and |log a| = 2; the perturbed 1D Sod state on two elements puts a pair there
  (1.2e-12 measured);
* the time loop started at t != t0 with a final step shorter than CFL*dt0 (dt = T - t,
  SSPRK33.jl:30): the same number of steps and the same final time, bit for bit;
* the mixed-precision batch on a single periodic element: rhs against the oracle within FP32
  term rounding, conservation and positivity over steps.
"""
import dataclasses

import numpy as np
import pytest

from conftest import normwise_rel, perturbed_state, problem, rel2
from oracle.oracle import Oracle
from test_gpu_mixed import TOL_TERMS, term_err, term_scale
from test_gpu_parity import TOL_INF

pytestmark = pytest.mark.gpu

CASES = [
    ("periodic_wave", dict(Kx=1, Ky=1, N=3, kx=1, ky=1, u0=0.3, v0=0.2)),
    # N = 3: at N = 2 the nodes of a one-element-wide strip sit on the zeros of sin(2 pi x) and
    # the state is uniform (free stream, rhs zero to rounding, no parity to measure)
    ("periodic_wave", dict(Kx=1, Ky=4, N=3, kx=1, ky=2, u0=0.3, v0=-0.2)),
    ("periodic_wave", dict(Kx=7, Ky=3, N=3, kx=2, ky=1, u0=0.5, v0=0.1, dip=(0.3, 0.6))),
    ("periodic_wave", dict(Kx=3, Ky=2, N=3, kx=1, ky=1, u0=0.5, v0=0.1, dip=(0.3, 0.6), basis="gauss")),
    ("sod_1d", dict(K=1, N=3)),
    ("sod_1d", dict(K=2, N=2)),
]


def _ids(c):
    kw = c[1]
    keys = ("Kx", "Ky", "K", "N", "basis")
    return c[0] + "-" + "-".join(f"{k}{kw[k]}" for k in keys if k in kw) + ("-dip" if "dip" in kw else "")


@pytest.mark.parametrize("case", CASES, ids=[_ids(c) for c in CASES])
@pytest.mark.parametrize("mode", [0, 1], ids=["ref-bounds", "intended-bounds"])
def test_edge_rhs_parity(case, mode):
    from p2de_jl_amd.solver import DeviceSolver, FIELD_RHSH, FIELD_RHSL
    name, kw = case
    solver, state, sp, fp = problem(name, **kw)
    fp = dataclasses.replace(fp, upper_bound_mode=mode)
    print(f"case {name} {kw} mode {mode}", flush=True)
    dev = DeviceSolver(fp)
    dev.set_debug(True)
    orc = Oracle(fp)
    U0 = state.preallocation.Uq
    for trial, U in enumerate([U0, perturbed_state(U0, seed=6)]):
        dev.set_state(U)
        for nstage, dt in [(1, 0.9 * fp.CFL * fp.dt0), (2, 0.5 * fp.CFL * fp.dt0), (3, 0.7 * fp.CFL * fp.dt0)]:
            dt_gpu, r_gpu = dev.rhs(0.0, dt, nstage)
            dt_cpu = orc.rhs(U, 0.0, dt, nstage)
            e = [normwise_rel(r_gpu, orc.field("rhsU"), fp.Nc),
                 normwise_rel(dev.field(FIELD_RHSH), orc.field("rhsH"), fp.Nc),
                 normwise_rel(dev.field(FIELD_RHSL), orc.field("rhsL"), fp.Nc)]
            e2 = rel2(r_gpu, orc.field("rhsU"), fp.Nc)
            print(f"  trial {trial} stage {nstage}: rhsU {e[0]:.2e} rhsH {e[1]:.2e} rhsL {e[2]:.2e} "
                  f"(2-norm {e2:.2e}) dt {dt_gpu:.17g} vs {dt_cpu:.17g}", flush=True)
            assert max(e) <= TOL_INF and e2 <= TOL_INF
            assert abs(dt_gpu - dt_cpu) <= 1e-14 * dt_cpu


@pytest.mark.parametrize("case", CASES, ids=[_ids(c) for c in CASES])
def test_edge_steps(case):
    from p2de_jl_amd.solver import DeviceSolver
    name, kw = case
    solver, state, sp, fp = problem(name, **kw)
    print(f"case {name} {kw}", flush=True)
    dev = DeviceSolver(fp)
    orc, twin = Oracle(fp), Oracle(fp)
    U = state.preallocation.Uq.copy()
    U2 = U * (1 + 1e-15 * np.random.default_rng(3).uniform(-1, 1, U.shape))
    dev.set_state(U)
    t_cpu = t_gpu = t2 = 0.0
    for step in range(3):
        t_cpu, dt_cpu = orc.step(U, t_cpu)
        t2, _ = twin.step(U2, t2)
        t_gpu, dt_gpu = dev.step(t_gpu)
        err = normwise_rel(dev.get_state(), U, fp.Nc)
        sens = normwise_rel(U2, U, fp.Nc)
        print(f"  step {step}: dt {dt_gpu:.17g} vs {dt_cpu:.17g} state err {err:.2e} (oracle sensitivity {sens:.1e})",
              flush=True)
        assert dt_gpu == dt_cpu and t_gpu == t_cpu
        assert err <= max(1e-12, 100 * sens)
        rep = dev.check_positivity()
        assert rep.violations == 0 and rep.min_rho > 0 and rep.min_p > 0


def test_late_start_and_short_last_step():
    """The loop from t = 0.3 (t0 = 0) to T = 0.3 + 2.37 CFL*dt0: two full steps and a short one
    of dt = T - t, then nothing (SSPRK33.jl:28-30); p2de_run against the oracle's loop."""
    from p2de_jl_amd.solver import DeviceSolver
    solver, state, sp, fp = problem("periodic_wave", Kx=6, Ky=5, N=3, kx=1, ky=1, u0=0.4, v0=0.3)
    fp = dataclasses.replace(fp, T=0.3 + 2.37 * fp.CFL * fp.dt0)
    dev = DeviceSolver(fp)
    orc = Oracle(fp)
    U = state.preallocation.Uq.copy()
    dev.set_state(U)
    dev.set_time(0.3)
    st = dev.run()
    t, n, dts = 0.3, 0, []
    while t < fp.T:
        t, dt = orc.step(U, t)
        dts.append(dt)
        n += 1
    t_gpu, last_dt = dev.get_time()
    print(f"  {st.steps} steps to t = {t_gpu!r} (last dt {last_dt!r}); oracle {n} steps to {t!r}, dts {dts}",
          flush=True)
    assert st.steps == n == 3 and t_gpu == t and last_dt == dts[-1]
    assert dts[-1] < 0.5 * fp.CFL * fp.dt0
    assert normwise_rel(dev.get_state(), U, fp.Nc) <= 1e-12
    assert dev.run().steps == 0


def test_mixed_single_element():
    from p2de_jl_amd.solver import FIELD_RHSH, FIELD_RHSL, BatchSolver
    solver, state, sp, fp = problem("periodic_wave", Kx=1, Ky=1, N=3, kx=1, ky=1, u0=0.3, v0=0.2)
    U0 = state.preallocation.Uq
    bs = BatchSolver([fp])
    bs.set_debug(True)
    orc = Oracle(fp)
    bs.set_state(0, U0)
    dt = 0.9 * fp.CFL * fp.dt0
    dt_gpu, r_gpu = bs.rhs(0, 0.0, dt, 1)
    dt_cpu = orc.rhs(U0, 0.0, dt, 1)
    sc = term_scale(fp, U0)
    eH = term_err(bs.field(0, FIELD_RHSH), orc.field("rhsH"), sc)
    eL = term_err(bs.field(0, FIELD_RHSL), orc.field("rhsL"), sc)
    print(f"  rhsH {eH:.2e} rhsL {eL:.2e} (term scale) dt {dt_gpu!r} vs {dt_cpu!r}", flush=True)
    assert max(eH, eL) <= TOL_TERMS and abs(dt_gpu - dt_cpu) <= 1e-6 * dt_cpu
    wJ = fp.wq[None, :] * fp.Jq.reshape(fp.K, fp.Nq)
    m0 = np.einsum("kq,kqc->c", wJ, U0)
    bs.set_state(0, U0)
    bs.set_time(0.0)
    bs.enqueue_steps(20)
    U = bs.get_state(0)
    m1 = np.einsum("kq,kqc->c", wJ, U)
    assert np.max(np.abs(m1 - m0) / np.abs(m0).clip(1e-300)) <= 1e-13
    rep = bs.check_positivity(0)
    assert rep.violations == 0 and rep.min_rho > 0 and rep.min_p > 0
