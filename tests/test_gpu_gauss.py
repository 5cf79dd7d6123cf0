"""GPU parity of the Gauss-collocation path (GaussCollocation with the NodewiseScaledExtrapolation
entropy-projection limiter, and without it) through the C-ABI against the CPU oracle.

Same bar as the Lobatto drop-in (tests/test_gpu_parity.py): per RHS evaluation and conserved
component, relative 2-norm error <= 1e-12 and worst-node error <= 1e-11 for rhsU, rhsH and
rhsL; dt to 1e-14; the limiter parameters L_local where they are determined to 1e-12.  The
entropy-projection parameters theta (bisection points, multiples of 2^-21) must agree
exactly: every bound test the bisection makes is compared with a margin far above the
one-ulp differences of the host's and the GPU's exp/log/pow.
"""
import numpy as np
import pytest

from conftest import normwise_rel, perturbed_state, problem, rel2
from oracle.oracle import Oracle
from test_gpu_parity import TOL, TOL_INF, limiter_mismatch

pytestmark = pytest.mark.gpu

CASES = [
    ("vortex_2d", dict(K=8, basis="gauss")),
    ("vortex_2d", dict(K=6, N=1, basis="gauss")),
    ("vortex_2d", dict(K=6, N=2, basis="gauss")),
    ("vortex_2d", dict(K=5, N=4, basis="gauss")),
    ("riemann4_2d", dict(K=12, basis="gauss")),
    ("kelvin_helmholtz_gauss", dict(K=8)),
    ("smoke_vortex", dict(N=3, basis="gauss")),
    ("smoke_vortex", dict(N=2, basis="gauss")),
    ("smoke_vortex", dict(N=3, basis="gauss_noep")),
    ("sod_1d", dict(basis="gauss")),
    ("sod_1d", dict(K=17, N=1, basis="gauss")),
    ("sod_1d", dict(K=9, N=4, basis="gauss")),
]


def _ids(c):
    return f"{c[0]}-" + "-".join(f"{k}{v}" for k, v in c[1].items())


def cold_node(U, k=14, i=5):
    """One node with p = 1e-6 and zero velocity: where its Lagrange weight at a face is
    negative the interpolated v4 turns positive and the limiter must act (theta = 0)."""
    U = U.copy()
    U[k, i, 1:-1] = 0.0
    U[k, i, -1] = 1e-6 / 0.4
    return U


def check_rhs(dev, orc, fp, U, nstage, dt):
    from p2de_jl_amd.solver import FIELD_L_LOCAL, FIELD_RHSH, FIELD_RHSL, FIELD_THETA
    dt_gpu, r_gpu = dev.rhs(0.0, dt, nstage)
    dt_cpu = orc.rhs(U, 0.0, dt, nstage)
    Nc = fp.Nc
    e_u = normwise_rel(r_gpu, orc.field("rhsU"), Nc)
    e_h = normwise_rel(dev.field(FIELD_RHSH), orc.field("rhsH"), Nc)
    e_l = normwise_rel(dev.field(FIELD_RHSL), orc.field("rhsL"), Nc)
    e2 = max(rel2(r_gpu, orc.field("rhsU"), Nc), rel2(dev.field(FIELD_RHSH), orc.field("rhsH"), Nc),
             rel2(dev.field(FIELD_RHSL), orc.field("rhsL"), Nc))
    e_L = limiter_mismatch(orc, dev.field(FIELD_L_LOCAL), fp, nstage)
    nlim, th_bad = 0, 0
    if fp.entropyproj_limiter:
        th_cpu = orc.field("theta_local").reshape(3, fp.K * fp.Nfp)[nstage - 1]
        th_gpu = dev.field(FIELD_THETA)
        nlim = int(np.sum(th_cpu < 1))
        th_bad = int(np.sum(th_gpu != th_cpu))
    print(f"  stage {nstage}: 2-norm {e2:.2e} inf: rhsU {e_u:.2e} rhsH {e_h:.2e} rhsL {e_l:.2e} L {e_L:.2e} "
          f"theta<1: {nlim} theta mismatches: {th_bad} dt {dt_gpu:.17g} vs {dt_cpu:.17g}", flush=True)
    assert th_bad == 0
    assert e2 <= TOL
    assert max(e_u, e_h, e_l) <= TOL_INF
    assert e_L <= 1e-12
    assert abs(dt_gpu - dt_cpu) <= 1e-14 * abs(dt_cpu)
    return nlim


@pytest.mark.parametrize("case", CASES, ids=[_ids(c) for c in CASES])
def test_gauss_rhs_parity(case):
    from p2de_jl_amd.solver import DeviceSolver
    name, kw = case
    solver, state, sp, fp = problem(name, **kw)
    print(f"case {name} {kw}", flush=True)
    dev = DeviceSolver(fp)
    dev.set_debug(True)
    orc = Oracle(fp)
    U0 = state.preallocation.Uq
    for trial, U in enumerate([U0, perturbed_state(U0, seed=1)]):
        dev.set_state(U)
        for nstage, dt in [(1, 0.9 * fp.CFL * fp.dt0), (2, 0.5 * fp.CFL * fp.dt0), (3, 0.7 * fp.CFL * fp.dt0)]:
            check_rhs(dev, orc, fp, U, nstage, dt)


@pytest.mark.parametrize("N", [2, 3, 4])
def test_gauss_limited_projection_parity(N):
    """A cold node forces theta < 1 (down to 0) on the faces of its lines."""
    from p2de_jl_amd.solver import DeviceSolver
    solver, state, sp, fp = problem("periodic_wave", Kx=6, Ky=6, N=N, basis="gauss")
    print(f"case periodic_wave 6x6 N={N} with a cold node", flush=True)
    U = cold_node(state.preallocation.Uq, i=N + 2)
    dev = DeviceSolver(fp)
    dev.set_debug(True)
    orc = Oracle(fp)
    dev.set_state(U)
    nlim = 0
    for nstage, dt in [(1, 1e-5), (2, 1e-5)]:
        nlim += check_rhs(dev, orc, fp, U, nstage, dt)
    assert nlim > 0


@pytest.mark.parametrize("case", [("smoke_vortex", dict(N=3, basis="gauss")), ("riemann4_2d", dict(K=12, basis="gauss")),
                                  ("sod_1d", dict(basis="gauss"))], ids=["smoke", "riemann4", "sod"])
def test_gauss_ssprk3_steps_parity(case):
    from p2de_jl_amd.solver import DeviceSolver
    name, kw = case
    solver, state, sp, fp = problem(name, **kw)
    print(f"case {name} {kw}", flush=True)
    dev = DeviceSolver(fp)
    orc = Oracle(fp)
    U = state.preallocation.Uq.copy()
    dev.set_state(U)
    t_cpu = t_gpu = 0.0
    nlim = 0
    for step in range(4):
        t_cpu, dt_cpu = orc.step(U, t_cpu)
        nlim += int(np.sum(orc.field("theta_local") < 1))
        t_gpu, dt_gpu = dev.step(t_gpu)
        Ug = dev.get_state()
        err = normwise_rel(Ug, U, fp.Nc)
        print(f"  step {step}: dt {dt_gpu:.17g} vs {dt_cpu:.17g} state err {err:.2e} theta<1 so far {nlim}", flush=True)
        assert abs(dt_gpu - dt_cpu) <= 1e-13 * dt_cpu
        assert err <= 1e-12
        rep = dev.check_positivity()
        assert rep.violations == 0 and rep.min_rho > 0 and rep.min_p > 0
    assert abs(t_gpu - t_cpu) <= 1e-14


@pytest.mark.parametrize("case", [("smoke_vortex", dict(N=3, basis="gauss")), ("vortex_2d", dict(K=5, N=4, basis="gauss")),
                                  ("vortex_2d", dict(K=6, N=2, basis="gauss")), ("sod_1d", dict(basis="gauss"))],
                         ids=["N3", "N4", "N2", "sod"])
def test_gauss_launch_variants_bit_identical(case):
    from p2de_jl_amd.solver import DeviceSolver
    name, kw = case
    solver, state, sp, fp = problem(name, **kw)
    print(f"case {name} {kw}", flush=True)
    dev = DeviceSolver(fp)
    U0 = state.preallocation.Uq
    out = []
    for v in range(dev.num_variants()):
        dev.set_variant(v)
        dev.set_state(U0)
        t, dt = dev.step(0.0)
        t, dt = dev.step(t)
        out.append(dev.get_state())
        print(f"  variant {dev.variant()}: t={t:.6e}", flush=True)
    for U in out[1:]:
        assert np.array_equal(U, out[0])
