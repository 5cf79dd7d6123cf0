"""End-to-end pins of the CPU oracle (CPU only): properties the restated scheme must have
whatever the arithmetic details, chosen so a dropped term, a wrong sign or a wrong index
fails one of them.

* free-stream preservation and discrete conservation of rhsH, rhsL and rhsU;
* the volume term's entropy identity  sum_h v_h . QF1_h = -sum_f B_f psi_f
  (experimental/check.jl:38-73), to the 1e-9 the reference's log-mean series allows;
* symmetry of the graph-viscosity coefficients and positivity of the low-order update
  under the CFL dt it returns;
* the limiter's convex-bound property: rho(U + dt rhsU) >= zeta rho(u_L) and
  rhoe(U + dt rhsU) >= zeta rhoe(u_L) at every node, u_L = U + dt rhsL, and symmetric
  parameters across element faces;
* one SSP33! step equals the Butcher form of SSP-RK3;
* accuracy: the vortex converges at high order with the intended bounds and at first
  order with the reference's; the Sod tube approaches the exact Riemann solution.
"""
import math

import numpy as np
import pytest

from conftest import PKG, normwise_rel, perturbed_state, problem
from oracle.oracle import Oracle

G = 1.4


def _pressure(U):
    Nd = U.shape[-1] - 2
    return (G - 1) * (U[..., -1] - 0.5 * np.sum(U[..., 1:1 + Nd] ** 2, axis=-1) / U[..., 0])


def _rhoe(U):
    Nd = U.shape[-1] - 2
    return U[..., -1] - 0.5 * np.sum(U[..., 1:1 + Nd] ** 2, axis=-1) / U[..., 0]


def _wJ(solver, fp):
    return (solver.md.J.T * fp.wq[None, :])        # (K, Nq)


@pytest.mark.parametrize("name,kw", [("vortex_2d", dict(K=6)), ("vortex_2d", dict(K=5, N=4)), ("sod_1d", dict(K=9))])
def test_free_stream_preservation(name, kw):
    solver, state, sp, fp = problem(name, **kw)
    if fp.Nd == 2:
        U = np.broadcast_to(np.array([1.2, 0.3, -0.4, 3.0]), state.preallocation.Uq.shape).copy()
        sp = PKG.StateParam(bcdata=sp.bcdata)
    else:
        U = np.broadcast_to(np.array([1.2, 0.3, 3.0]), state.preallocation.Uq.shape).copy()
        # periodic 1D mesh for a free stream
        md = PKG.make_periodic(solver.md)
        sp = PKG.StateParam(bcdata=PKG.operators.BCData(mapP=md.mapP))
        fp = PKG.flat_problem(solver, sp)
    o = Oracle(fp)
    o.rhs(U, 0.0, 1e-3, 1)
    for f in ("rhsH", "rhsL", "rhsU"):
        assert np.max(np.abs(o.field(f))) <= 1e-12, f


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("name,kw", [("vortex_2d", dict(K=6)), ("kelvin_helmholtz", dict(K=6)), ("smoke_vortex", dict(N=2))])
def test_discrete_conservation(name, kw, mode):
    solver, state, sp, fp = problem(name, **kw)
    o = Oracle(fp, upper_bound_mode=mode)
    U = perturbed_state(state.preallocation.Uq, seed=4)
    o.rhs(U, 0.0, 0.5 * fp.CFL * fp.dt0, 1)
    wJ = _wJ(solver, fp)
    for f in ("rhsH", "rhsL", "rhsU"):
        r = o.field(f).reshape(fp.K, fp.Nq, fp.Nc)
        tot = np.sum(wJ[:, :, None] * r, axis=(0, 1))
        scale = np.sum(np.abs(wJ[:, :, None] * r), axis=(0, 1))
        assert np.all(np.abs(tot) <= 1e-13 * scale), (f, tot, scale)


@pytest.mark.parametrize("name,kw", [("vortex_2d", dict(K=6)), ("kelvin_helmholtz", dict(K=6, N=3)), ("sod_1d", dict(K=12))])
def test_volume_term_entropy_identity(name, kw):
    """sum_h v_h . QF1_h + sum_f B_f psi_f = 0 per element and direction, with v_h the
    entropy variables the fluxes were formed from (check.jl:38-73)."""
    solver, state, sp, fp = problem(name, **kw)
    o = Oracle(fp)
    U = perturbed_state(state.preallocation.Uq, seed=2, amp=0.2)
    o.rhs(U, 0.0, 1e-4, 1)
    K, Nq, Nfp, Nh, Nc, Nd = fp.K, fp.Nq, fp.Nfp, fp.Nh, fp.Nc, fp.Nd
    vq = o.field("vq").reshape(K, Nq, Nc)
    vh = np.concatenate([vq, vq[:, fp.fq2q - 1, :]], axis=1)     # LGL: v_tilde_f = vq[fq2q]
    QF1 = o.field("QF1").reshape(K, Nh, Nd, Nc)
    psi = o.field("psi_tilde").reshape(K, Nh, Nd)
    ops = solver.discrete_data.ops
    rxJ = solver.md.GJh[0][0, 0]
    syJ = solver.md.GJh[-1][0, 0] if Nd == 2 else None
    for d in range(Nd):
        B = np.diag(ops.Brs[d]) * (rxJ if d == 0 else syJ)
        lhs = np.einsum("khc,khc->k", vh, QF1[:, :, d, :])
        rhs = -np.einsum("f,kf->k", B, psi[:, Nq:, d])
        scale = np.einsum("khc,khc->k", np.abs(vh), np.abs(QF1[:, :, d, :])) + np.abs(rhs)
        assert np.all(np.abs(lhs - rhs) <= 2e-9 * scale), np.max(np.abs(lhs - rhs) / scale)


@pytest.mark.parametrize("name,kw", [("riemann4_2d", dict(K=12)), ("sod_1d", dict(K=30)), ("vortex_2d", dict(K=6))])
def test_low_order_update_is_positive(name, kw):
    solver, state, sp, fp = problem(name, **kw)
    fp.dt0 = 1e30                                   # let the positivity CFL bind
    o = Oracle(fp)
    for seed in range(3):
        U = perturbed_state(state.preallocation.Uq, seed=seed, amp=0.3)
        dt = o.rhs(U, 0.0, 1e-4, 1)
        lam = o.field("lambda").reshape(fp.K, fp.Nq, fp.Nq)
        assert np.array_equal(lam, np.transpose(lam, (0, 2, 1)))
        W = U + dt * o.field("rhsL").reshape(U.shape)
        assert np.all(W[..., 0] > 0) and np.all(_pressure(W) > 0)


def _symmetric_across_faces(o, fp, nstage):
    NLs = fp.Nq + fp.N1D
    L = o.field("L_local").reshape(3, fp.K, fp.Nd, NLs)[nstage - 1]
    N1D, Nfp = fp.N1D, fp.Nfp
    mapP = fp.mapP.reshape(fp.K, Nfp) - 1
    for k in range(fp.K):
        for f in range(Nfp):
            kP, iP = divmod(int(mapP[k, f]), Nfp)
            def idx(face):
                if face < N1D:
                    return 0, (N1D + 1) * face
                if face < 2 * N1D:
                    return 0, N1D + (N1D + 1) * (face - N1D)
                if face < 3 * N1D:
                    return 1, face - 2 * N1D
                return 1, (face - 3 * N1D) + N1D * N1D
            d, a = idx(f)
            dP, b = idx(iP)
            assert L[k, d, a] == L[kP, dP, b]


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("name,kw", [("riemann4_2d", dict(K=10)), ("vortex_2d", dict(K=6)), ("sod_1d", dict(K=30)),
                                     ("kelvin_helmholtz", dict(K=6))])
def test_limiter_convex_bound_property(name, kw, mode):
    solver, state, sp, fp = problem(name, **kw)
    o = Oracle(fp, upper_bound_mode=mode)
    zeta = fp.zeta
    for seed in range(2):
        U = perturbed_state(state.preallocation.Uq, seed=seed, amp=0.3)
        dtpos = o.rhs(U, 0.0, 1e-4, 1)
        dt = min(dtpos, fp.CFL * fp.dt0)
        o.rhs(U, 0.0, dt, 2)                        # stage 2: limiter with this dt
        uL = U + dt * o.field("rhsL").reshape(U.shape)
        W = U + dt * o.field("rhsU").reshape(U.shape)
        L = o.field("L_local").reshape(3, -1)[1]
        assert np.all((L >= 0) & (L <= 1))
        assert np.all(W[..., 0] >= zeta * uL[..., 0] * (1 - 1e-12))
        assert np.all(_rhoe(W) >= zeta * _rhoe(uL) - 1e-12 * np.abs(_rhoe(uL)))
        if fp.Nd == 2:
            _symmetric_across_faces(o, fp, 2)


def test_limiter_modes_reduce_to_low_and_high_order():
    """Periodic smooth flow: the reference's bounds give the low-order RHS (l = 0 on every
    face whose density flux jump is nonzero); the intended bounds leave the high-order RHS."""
    solver, state, sp, fp = problem("vortex_2d", K=8)
    U = state.preallocation.Uq
    dt = 0.9 * fp.CFL * fp.dt0
    o0 = Oracle(fp, upper_bound_mode=0)
    o0.rhs(U, 0.0, dt, 1)
    assert normwise_rel(o0.field("rhsU"), o0.field("rhsL"), fp.Nc) < 1e-12
    o1 = Oracle(fp, upper_bound_mode=1)
    o1.rhs(U, 0.0, dt, 1)
    assert normwise_rel(o1.field("rhsU"), o1.field("rhsH"), fp.Nc) < 1e-12


@pytest.mark.parametrize("name,kw", [("vortex_2d", dict(K=6)), ("sod_1d", dict(K=20)), ("riemann4_2d", dict(K=8))])
def test_ssprk3_step_equals_butcher_form(name, kw):
    solver, state, sp, fp = problem(name, **kw)
    o = Oracle(fp)
    U = perturbed_state(state.preallocation.Uq, seed=5, amp=0.05)
    t = 0.0
    dt_in = min(fp.CFL * fp.dt0, fp.T - t)
    dt = o.rhs(U, t, dt_in, 1)
    k1 = o.field("rhsU").reshape(U.shape)
    o.rhs(U + dt * k1, t, dt, 2)
    k2 = o.field("rhsU").reshape(U.shape)
    o.rhs(U + dt / 4 * (k1 + k2), t, dt, 3)
    k3 = o.field("rhsU").reshape(U.shape)
    ref = U + dt / 6 * (k1 + k2 + 4 * k3)
    V = U.copy()
    t1, dt1 = o.step(V, t)
    assert dt1 == dt and t1 == t + dt
    assert normwise_rel(V, ref, fp.Nc) <= 1e-13


def _vortex_error(K, N, mode, T=0.2):
    # dt0 from the positivity dt of the initial state: the stage-1 limiter uses
    # min(CFL*dt0, T - t) (rhs.jl:46,52), so dt0 must not be left unbounded
    solver, state, sp, fp = problem("vortex_2d", K=K, N=N, T=T, dt0=1e30)
    dt_pos = Oracle(fp).rhs(state.preallocation.Uq, 0.0, 1e-6, 1)
    solver, state, sp, fp = problem("vortex_2d", K=K, N=N, T=T, dt0=0.9 * dt_pos / 0.5)
    o = Oracle(fp, upper_bound_mode=mode)
    U = state.preallocation.Uq.copy()
    t = 0.0
    while t < T:
        t, dt = o.step(U, t)
    rho_ex = PKG.configs.vortex_exact_prim(solver.md.xq, solver.md.yq, T)[0].T
    wJ = _wJ(solver, fp)
    return np.sum(wJ * np.abs(U[..., 0] - rho_ex)) / np.sum(wJ * np.abs(rho_ex))


def test_vortex_convergence_rates():
    e8, e16 = _vortex_error(8, 3, 1), _vortex_error(16, 3, 1)
    assert e8 / e16 > 6.0, (e8, e16)                   # high order with the intended bounds
    r8, r16 = _vortex_error(8, 3, 0), _vortex_error(16, 3, 0)
    assert 1.3 < r8 / r16 < 3.0, (r8, r16)             # first order with the reference's bounds
    assert e16 < r16


def exact_riemann_density(x, t, WL=(1.0, 0.0, 1.0), WR=(0.125, 0.0, 0.1), x0=0.5, g=1.4):
    """Exact solution of the 1D Riemann problem (Toro, ch. 4): density at x, t."""
    rL, uL, pL = WL
    rR, uR, pR = WR
    cL, cR = math.sqrt(g * pL / rL), math.sqrt(g * pR / rR)

    def fK(p, r, pk, ck):
        if p > pk:
            A, B = 2 / ((g + 1) * r), (g - 1) / (g + 1) * pk
            return (p - pk) * math.sqrt(A / (p + B))
        return 2 * ck / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1)

    lo, hi = 1e-12, 10.0
    for _ in range(200):
        pm = 0.5 * (lo + hi)
        if fK(pm, rL, pL, cL) + fK(pm, rR, pR, cR) + (uR - uL) > 0:
            hi = pm
        else:
            lo = pm
    ps = 0.5 * (lo + hi)
    us = 0.5 * (uL + uR) + 0.5 * (fK(ps, rR, pR, cR) - fK(ps, rL, pL, cL))
    out = np.empty_like(x)
    for n, xx in enumerate(x.ravel()):
        s = (xx - x0) / t
        if s <= us:
            if ps <= pL:   # left rarefaction
                shl, stl = uL - cL, us - cL * (ps / pL) ** ((g - 1) / (2 * g))
                if s <= shl:
                    r = rL
                elif s >= stl:
                    r = rL * (ps / pL) ** (1 / g)
                else:
                    r = rL * (2 / (g + 1) + (g - 1) / ((g + 1) * cL) * (uL - s)) ** (2 / (g - 1))
            else:
                SL = uL - cL * math.sqrt((g + 1) / (2 * g) * ps / pL + (g - 1) / (2 * g))
                r = rL if s <= SL else rL * ((ps / pL + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pL + 1))
        else:
            if ps > pR:    # right shock
                SR = uR + cR * math.sqrt((g + 1) / (2 * g) * ps / pR + (g - 1) / (2 * g))
                r = rR if s >= SR else rR * ((ps / pR + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pR + 1))
            else:
                shr, str_ = uR + cR, us + cR * (ps / pR) ** ((g - 1) / (2 * g))
                if s >= shr:
                    r = rR
                elif s <= str_:
                    r = rR * (ps / pR) ** (1 / g)
                else:
                    r = rR * (2 / (g + 1) - (g - 1) / ((g + 1) * cR) * (uR - s)) ** (2 / (g - 1))
        out.ravel()[n] = r
    return out


def test_exact_riemann_solver_star_state():
    # Toro's table 4.2, test 1 (Sod): p* = 0.30313, u* = 0.92745
    x = np.array([0.5 + 0.92745 * 0.2 - 1e-6])
    assert abs(exact_riemann_density(x, 0.2)[0] - 0.42632) < 2e-4


@pytest.mark.parametrize("mode", [0, 1])
def test_sod_l1_error_against_exact_solution(mode):
    errs = []
    for K in (50, 100):
        solver, state, sp, fp = problem("sod_1d", K=K)
        o = Oracle(fp, upper_bound_mode=mode)
        U = state.preallocation.Uq.copy()
        t = 0.0
        while t < fp.T:
            t, dt = o.step(U, t)
            assert o.count_nonpositive(U)[0] == 0
        rho_ex = exact_riemann_density(solver.md.xq, fp.T)
        wJ = solver.md.J * fp.wq[:, None]
        errs.append(np.sum(wJ * np.abs(U[..., 0].T - rho_ex)))
    assert errs[1] < 0.02 and errs[1] < errs[0]
