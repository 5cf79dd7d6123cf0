"""The oracle's Gauss-collocation path (CPU): what pins it.

The reference has no Gauss test or fixture (test/test_smoke.jl runs Lobatto only) and its
operators come from StartUpDG, which is absent here, so the Gauss branch of the oracle
(entropy projection with NodewiseScaledExtrapolation, filter.jl:6-130, rhs.jl:113-133,
flux_differencing.jl:288-328) is pinned by what the mathematics fixes:

* rhsH of a smooth translating vortex converges to the exact time derivative at order ~N
  (a dropped term, a wrong sign or a transposed operand leaves an O(1) error);
* the limited projection keeps every line's subcell flux potentials telescoping onto the far
  face flux (rows of Vf_new sum to one and stay on their line): a transposed Vf_new fails it;
* a cold node that makes the unlimited projection of the entropy variables ill-defined (NaN
  face states) is repaired by the limiter (theta = 0 there, every face state admissible);
* theta is a dyadic bisection point in [0, 1]; with theta = 1 everywhere the limited and
  unlimited operators give the same bits;
* the discrete entropy identity of the hybridized scheme holds on every element, limited or
  not: sum_i wJ v_i . rhsH_i = sum_f (B_f . psi~_f - v~_f . BF_H,f).  It needs the assembly
  (MinvVhT_new) and the projection of the face states to use the same Vf_new, so a blend
  written differently in either place fails it (checked by planting one);
* periodic problems conserve every component through limited steps.
"""
import numpy as np
import pytest

from conftest import PKG
from oracle.oracle import Oracle, pt_v_of_u

C = PKG.configs


def _setup(name, basis, **kw):
    s, st, sp = C.build(name, basis=basis, **kw)
    return s, st.preallocation.Uq.copy(), PKG.flat_problem(s, sp)


def _vortex_dUdt(s, d=1e-4):
    x, y = s.md.xq, s.md.yq
    up = C.primitive_to_conservative(1.4, C.vortex_exact_prim(x, y, d))
    um = C.primitive_to_conservative(1.4, C.vortex_exact_prim(x, y, -d))
    return np.transpose((up - um) / (2 * d), (2, 1, 0))


@pytest.mark.parametrize("basis", ["gauss", "lobatto"])
@pytest.mark.parametrize("N,tol", [(3, 1e-2), (4, 2e-3)])
def test_high_order_rhs_converges_to_the_exact_time_derivative(basis, N, tol):
    errs = []
    for K in (16, 32):
        s, U, fp = _setup("vortex_2d", basis, K=K, N=N)
        orc = Oracle(fp)
        orc.rhs(U, 0.0, 1e-6, 2)
        rH = orc.field("rhsH").reshape(fp.K, fp.Nq, 4)
        ex = _vortex_dUdt(s)
        wJ = (fp.wq[None, :] * fp.Jq.reshape(fp.K, fp.Nq))[..., None]
        errs.append(np.sqrt(np.sum(wJ * (rH - ex) ** 2) / np.sum(wJ * ex ** 2)))
    rate = np.log2(errs[0] / errs[1])
    assert errs[1] < tol and rate > 2.5, (errs, rate)


def _limited_state():
    """A periodic wave with one cold node (p = 1e-6) at a node whose Lagrange weight at the
    faces is negative: the interpolated v4 turns positive there without the limiter."""
    s, U, fp = _setup("periodic_wave", "gauss", Kx=6, Ky=6, N=3)
    k, i = 14, 5
    U[k, i, 1:3] = 0.0
    U[k, i, 3] = 1e-6 / 0.4
    return s, U, fp


def _face_states(orc, fp):
    ut = orc.field("u_tilde").reshape(fp.K, fp.Nh, fp.Nc)[:, fp.Nq:, :]
    rho = ut[..., 0]
    p = 0.4 * (ut[..., 3] - 0.5 * (ut[..., 1] ** 2 + ut[..., 2] ** 2) / rho)
    return np.isfinite(rho) & np.isfinite(p) & (rho > 0) & (p > 0)


def test_limiter_repairs_ill_defined_face_projection():
    s, U, fp = _limited_state()
    orc = Oracle(fp)
    orc.rhs(U, 0.0, 1e-5, 1)
    th = orc.field("theta_local").reshape(3, fp.K, fp.Nfp)[0]
    assert np.all(_face_states(orc, fp))
    assert np.sum(th < 1) >= 1 and th.min() == 0.0
    fp.entropyproj_limiter = 0                                  # GaussCollocation, no limiter
    orc0 = Oracle(fp)
    orc0.rhs(U, 0.0, 1e-5, 1)
    bad = ~_face_states(orc0, fp)
    assert bad.sum() >= 1 and np.array_equal(np.argwhere(bad), np.argwhere(th < 1))


def test_theta_is_a_dyadic_bisection_point():
    s, U, fp = _setup("smoke_vortex", "gauss")
    orc = Oracle(fp)
    t = 0.0
    seen = 0
    for _ in range(2):
        t, _ = orc.step(U, t)
        th = orc.field("theta_local")
        assert np.all((th >= 0) & (th <= 1))
        lim = th[th < 1]
        seen += lim.size
        assert np.array_equal(lim * 2.0 ** 21, np.round(lim * 2.0 ** 21))
    assert seen > 0


def _fbar_end_and_far_face(orc, fp):
    """(f_bar_H at s = N1D, BF_H of the far face, scale) per element, line and direction."""
    K, N1D, Nc, Nfp = fp.K, fp.N1D, fp.Nc, fp.Nfp
    NLs = fp.Nq + N1D
    fb = orc.field("f_bar_H").reshape(2, K, NLs, Nc)
    BF = orc.field("BF_H").reshape(K, Nfp, 2, Nc)
    ends, far, scale = [], [], []
    for line in range(N1D):
        # x lines: f_bar index si + (N1D+1)*sj, far face N1D + sj
        ends.append(fb[0][:, N1D + (N1D + 1) * line, :])
        far.append(BF[:, N1D + line, 0, :])
        # y lines: index si + N1D*sj, far face 3*N1D + si
        ends.append(fb[1][:, line + N1D * N1D, :])
        far.append(BF[:, 3 * N1D + line, 1, :])
    ends, far = np.stack(ends, 1), np.stack(far, 1)
    scale = np.abs(fb).max() + np.abs(BF).max()
    return ends, far, scale


def test_limited_projection_keeps_line_telescoping():
    s, U, fp = _limited_state()
    orc = Oracle(fp)
    orc.rhs(U, 0.0, 1e-5, 1)
    th = orc.field("theta_local").reshape(3, fp.K, fp.Nfp)[0]
    limited = th.min(axis=1) < 1
    assert limited.any()
    ends, far, scale = _fbar_end_and_far_face(orc, fp)
    err = np.abs(ends + far)
    assert err[limited].max() <= 1e-12 * scale and err.max() <= 1e-12 * scale


def test_theta_one_gives_the_unlimited_bits():
    s, U, fp = _setup("vortex_2d", "gauss", K=6)
    orc = Oracle(fp)
    orc.rhs(U, 0.0, 1e-3, 1)
    assert np.all(orc.field("theta_local")[: fp.K * fp.Nfp] == 1.0)
    r1 = orc.field("rhsU")
    fp.entropyproj_limiter = 0
    orc0 = Oracle(fp)
    orc0.rhs(U, 0.0, 1e-3, 1)
    assert np.array_equal(r1, orc0.field("rhsU"))


def test_limited_steps_conserve_and_stay_positive():
    s, U, fp = _setup("smoke_vortex", "gauss")
    wJ = fp.wq[None, :] * fp.Jq.reshape(fp.K, fp.Nq)
    m0 = np.einsum("kq,kqc->c", wJ, U)
    a0 = np.einsum("kq,kqc->c", wJ, np.abs(U))
    orc = Oracle(fp)
    t = 0.0
    nlim = 0
    for _ in range(2):
        t, _ = orc.step(U, t)
        nlim += int(np.sum(orc.field("theta_local") < 1))
    m1 = np.einsum("kq,kqc->c", wJ, U)
    assert nlim > 0
    assert np.max(np.abs(m1 - m0) / a0) < 1e-13
    assert orc.count_nonpositive(U)[0] == 0


def test_entropy_identity_on_limited_and_unlimited_elements():
    s, U, fp = _limited_state()
    orc = Oracle(fp)
    orc.rhs(U, 0.0, 1e-5, 1)
    K, Nq, Nfp, Nh, Nc = fp.K, fp.Nq, fp.Nfp, fp.Nh, fp.Nc
    th = orc.field("theta_local").reshape(3, K, Nfp)[0]
    assert (th.min(axis=1) < 1).any()
    rH = orc.field("rhsH").reshape(K, Nq, Nc)
    vq = orc.field("vq").reshape(K, Nq, Nc)
    ut = orc.field("u_tilde").reshape(K, Nh, Nc)[:, Nq:, :]
    vt = np.array([[pt_v_of_u(2, fp.gamma, ut[k, f]) for f in range(Nfp)] for k in range(K)])
    psi = orc.field("psi_tilde").reshape(K, Nh, 2)[:, Nq:, :]
    BF = orc.field("BF_H").reshape(K, Nfp, 2, Nc)
    Brs = fp.Brs.reshape(2, Nfp, Nfp)
    Br, Bs = np.diag(Brs[0]), np.diag(Brs[1])
    GJh = fp.GJh.reshape(4, K, Nh)[:, :, Nq:]
    Bx = GJh[0] * Br[None] + GJh[1] * Bs[None]
    By = GJh[2] * Br[None] + GJh[3] * Bs[None]
    wJ = fp.wq[None, :] * fp.Jq.reshape(K, Nq)
    lhs = np.einsum("kq,kqc,kqc->k", wJ, vq, rH)
    rhs = np.sum(Bx * psi[..., 0] + By * psi[..., 1], 1) - np.einsum("kfc,kfdc->k", vt, BF)
    scale = (np.einsum("kq,kqc,kqc->k", wJ, np.abs(vq), np.abs(rH))
             + np.einsum("kfc,kfdc->k", np.abs(vt), np.abs(BF)) + 1e-300)
    assert np.max(np.abs(lhs - rhs) / scale) < 1e-9
