"""Set-up operators and mesh (CPU): closed forms and identities that pin
``operators.py`` (the restatement of init.jl:133-340 and StartUpDG's element data)."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from conftest import PKG

ops_mod = PKG.operators
GOLD = json.loads((Path(__file__).parent / "golden" / "pointwise.json").read_text())


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_lgl_nodes_weights_closed_form_and_mpmath(N):
    x, w = ops_mod.lgl_nodes_weights(N)
    closed = {1: ([-1, 1], [1, 1]),
              2: ([-1, 0, 1], [1 / 3, 4 / 3, 1 / 3]),
              3: ([-1, -1 / math.sqrt(5), 1 / math.sqrt(5), 1], [1 / 6, 5 / 6, 5 / 6, 1 / 6]),
              4: ([-1, -math.sqrt(3 / 7), 0, math.sqrt(3 / 7), 1], [1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10])}
    assert np.allclose(x, closed[N][0], rtol=0, atol=2e-16)
    assert np.allclose(w, closed[N][1], rtol=0, atol=4e-16)
    g = GOLD["lgl"][str(N)]
    assert np.allclose(x, g["x"], rtol=0, atol=2e-16) and np.allclose(w, g["w"], rtol=0, atol=4e-16)


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_differentiation_exact_on_polynomials_and_sbp(N):
    x, w = ops_mod.lgl_nodes_weights(N)
    D = ops_mod.lagrange_diff_matrix(x)
    for p in range(N + 1):
        assert np.allclose(D @ x ** p, p * x ** max(p - 1, 0) if p else 0 * x, atol=1e-13)
    Q = np.diag(w) @ D
    B = np.zeros((N + 1, N + 1))
    B[0, 0], B[-1, -1] = -1.0, 1.0
    assert np.allclose(Q + Q.T, B, atol=1e-14)          # SBP property
    assert np.allclose(Q @ np.ones(N + 1), 0, atol=1e-14)


@pytest.mark.parametrize("Nd,N", [(1, 1), (1, 3), (1, 4), (2, 1), (2, 2), (2, 3), (2, 4)])
def test_hybridized_operators(Nd, N):
    rd = ops_mod.lobatto_ref_elem(N, Nd)
    ops, sz = ops_mod.build_operators(rd, 5e-16)
    N1D = N + 1
    Nq, Nfp, Nh = sz.Nq, sz.Nfp, sz.Nh
    assert Nq == N1D ** Nd and Nfp == (2 if Nd == 1 else 4 * N1D) and Nh == Nq + Nfp
    # pair counts (volume-volume within lines plus one face-volume pair per face node)
    nph = N1D * (N1D - 1) // 2 + 2 if Nd == 1 else N1D * N1D * (N1D - 1) + Nfp
    npl = N1D - 1 if Nd == 1 else 2 * N1D * (N1D - 1)
    assert len(ops.Srsh_nnz) == nph and len(ops.Srs0_nnz) == npl
    for d in range(Nd):
        S = ops.Srsh_db[d]
        assert np.allclose(S, -S.T, atol=0)                            # antisymmetric
        B = np.diag(ops.Brs[d])
        # Srsh_db 1 = -[0; diag(B)]: the hybridized SBP identity Qh 1 = 0, Qh + Qh^T = diag(0, B)
        assert np.allclose(S @ np.ones(Nh), -np.concatenate([np.zeros(Nq), B]), atol=1e-13)
        S0 = ops.Srs0[d]
        assert np.allclose(S0, -S0.T, atol=0)
        # Q0 1 = 0 and Q0 + Q0^T = B0 give S0 1 = -B0 1 / 2, with B0 = kron(M1D, B1D) for r
        B1 = np.zeros(N1D)
        B1[0], B1[-1] = -1.0, 1.0
        W1 = rd.w1D
        B0 = B1 if Nd == 1 else (np.kron(W1, B1) if d == 0 else np.kron(B1, W1))
        assert np.allclose(S0 @ np.ones(Nq), -0.5 * B0, atol=1e-15)
    # nnz lists are in the reference's order: column j ascending, then row i ascending
    for lst in (ops.Srsh_nnz, ops.Srs0_nnz):
        keys = [(j, i) for i, j in lst]
        assert keys == sorted(keys) and all(i > j for i, j in lst)
    # LGL: faces pass through nodes; Vf is a selection; MinvVhT = diag(1/w) [I Vf^T]
    assert np.array_equal(ops.Vf, ops.Vf_low)
    assert np.allclose(ops.MinvVhT, np.diag(1 / ops.wq) @ np.hstack([np.eye(Nq), ops.Vf.T]), rtol=1e-15, atol=0)
    for f, q in enumerate(ops.fq2q):
        assert ops.Vf[f, q - 1] == 1.0 and (f + 1) in ops.q2fq[q - 1]
    if Nd == 2:
        for m in range(N1D):   # face layout (subcell.jl:173-206)
            assert ops.fq2q[m] - 1 == m * N1D and ops.fq2q[N1D + m] - 1 == m * N1D + N1D - 1
            assert ops.fq2q[2 * N1D + m] - 1 == m and ops.fq2q[3 * N1D + m] - 1 == (N1D - 1) * N1D + m
        corners = [q for q in range(Nq) if len(ops.q2fq[q]) == 2]
        assert len(corners) == 4 and all(ops.q2fq[q][0] <= 2 * N1D < ops.q2fq[q][1] for q in corners)


@pytest.mark.parametrize("Nd", [1, 2])
def test_mesh_connectivity(Nd):
    N = 3
    rd = ops_mod.lobatto_ref_elem(N, Nd)
    ops, sz = ops_mod.build_operators(rd, 5e-16)
    if Nd == 1:
        md = ops_mod.uniform_mesh(rd, 7, 0.0, 1.0, periodic=True)
    else:
        md = ops_mod.uniform_mesh(rd, (5, 3), (0.0, -1.0), (2.0, 1.0), periodic=(True, True))
    K = md.J.shape[1]
    mapP = md.mapP - 1
    Nfp = mapP.shape[0]
    flat = mapP.T.ravel()
    assert np.array_equal(flat[flat], np.arange(Nfp * K))        # involution
    fq = ops.fq2q - 1
    xf = md.xq[fq, :]
    xP = xf.T.ravel()[flat].reshape(K, Nfp).T
    L = 1.0 if Nd == 1 else 2.0
    d = np.abs(xf - xP)
    assert np.all(np.minimum(d, np.abs(d - L)) < 1e-12)           # matching coordinates (mod period)
    if Nd == 2:
        yf = md.yq[fq, :]
        yP = yf.T.ravel()[flat].reshape(K, Nfp).T
        dy = np.abs(yf - yP)
        assert np.all(np.minimum(dy, np.abs(dy - 2.0)) < 1e-12)
        assert np.allclose(md.J, (2.0 / 5 / 2) * (2.0 / 3 / 2))
    # non-periodic: boundary face nodes map to themselves
    md2 = ops_mod.uniform_mesh(rd, 7 if Nd == 1 else (4, 4), 0.0 if Nd == 1 else (0.0, 0.0),
                               1.0 if Nd == 1 else (1.0, 1.0), periodic=False)
    m2 = (md2.mapP - 1).T.ravel()
    b = md2.boundary.T.ravel()
    assert np.array_equal(m2[b], np.flatnonzero(b))
    assert b.sum() == (2 if Nd == 1 else 4 * 4 * (N + 1))


def test_sod_interface_nodes_take_the_right_state():
    solver, state, sp = PKG.configs.build("sod_1d")
    x = solver.md.xq
    rho = state.preallocation.Uq[:, :, 0].T
    at_half = np.abs(x - 0.5) < 1e-15
    assert at_half.sum() == 2 and np.all(rho[at_half] == 0.125)


def test_unsupported_options_are_rejected():
    T = PKG.types
    base = PKG.configs.vortex_2d(K=4)[0]
    from dataclasses import replace
    base1 = PKG.configs.sod_1d(K=4)[0]
    for bad in (replace(base1, rhs_limiter=T.SubcellLimiter(bound=T.PositivityAndCellEntropyBound())),
                replace(base1, rhs_limiter=T.SubcellLimiter(bound=T.TVDAndCellEntropyBound())),
                replace(base, entropyproj_limiter=T.NodewiseScaledExtrapolation()),
                replace(base, N=5)):
        with pytest.raises(T.UnsupportedOption):
            T.check_supported(bad)


# --------------------------------------------------------------------------- Gauss basis
GAUSS_CLOSED = {
    1: ([-1 / math.sqrt(3), 1 / math.sqrt(3)], [1.0, 1.0]),
    2: ([-math.sqrt(3 / 5), 0.0, math.sqrt(3 / 5)], [5 / 9, 8 / 9, 5 / 9]),
    3: ([-math.sqrt(3 / 7 + 2 / 7 * math.sqrt(6 / 5)), -math.sqrt(3 / 7 - 2 / 7 * math.sqrt(6 / 5)),
         math.sqrt(3 / 7 - 2 / 7 * math.sqrt(6 / 5)), math.sqrt(3 / 7 + 2 / 7 * math.sqrt(6 / 5))],
        [(18 - math.sqrt(30)) / 36, (18 + math.sqrt(30)) / 36, (18 + math.sqrt(30)) / 36, (18 - math.sqrt(30)) / 36]),
    4: ([-1 / 3 * math.sqrt(5 + 2 * math.sqrt(10 / 7)), -1 / 3 * math.sqrt(5 - 2 * math.sqrt(10 / 7)), 0.0,
         1 / 3 * math.sqrt(5 - 2 * math.sqrt(10 / 7)), 1 / 3 * math.sqrt(5 + 2 * math.sqrt(10 / 7))],
        [(322 - 13 * math.sqrt(70)) / 900, (322 + 13 * math.sqrt(70)) / 900, 128 / 225,
         (322 + 13 * math.sqrt(70)) / 900, (322 - 13 * math.sqrt(70)) / 900]),
}


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_gauss_nodes_weights_closed_form(N):
    x, w = ops_mod.gauss_nodes_weights(N)
    assert np.allclose(x, GAUSS_CLOSED[N][0], rtol=0, atol=4e-16)
    assert np.allclose(w, GAUSS_CLOSED[N][1], rtol=0, atol=4e-16)
    assert np.array_equal(x, -x[::-1]) and np.array_equal(w, w[::-1])
    # exact for polynomials of degree 2N+1
    for p in range(2 * N + 2):
        assert abs(np.dot(w, x ** p) - (0.0 if p % 2 else 2.0 / (p + 1))) < 1e-14


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_interpolation_rows_exact_on_polynomials(N):
    x, _ = ops_mod.gauss_nodes_weights(N)
    for xi in (-1.0, 1.0, 0.3):
        row = ops_mod.lagrange_interp_row(x, xi)
        for p in range(N + 1):
            assert abs(row @ x ** p - xi ** p) < 1e-14


@pytest.mark.parametrize("Nd,N", [(1, 1), (1, 3), (1, 4), (2, 1), (2, 2), (2, 3), (2, 4)])
def test_gauss_hybridized_operators(Nd, N):
    """The Gauss element of init.jl:66-131 through build_operators: pair structure, the
    hybridized SBP identities, the SBP property Q + Q^T = E^T B E of the Gauss collocation
    operator, Vf exact on polynomials, Vf_low picking the line ends (init.jl:324-336)."""
    rd = ops_mod.gauss_ref_elem(N, Nd)
    ops, sz = ops_mod.build_operators(rd, 5e-16, quad="gauss")
    N1D = N + 1
    Nq, Nfp, Nh = sz.Nq, sz.Nfp, sz.Nh
    nph = N1D * (N1D - 1) // 2 + 2 * N1D if Nd == 1 else N1D * N1D * (N1D - 1) + Nfp * N1D
    npl = N1D - 1 if Nd == 1 else 2 * N1D * (N1D - 1)
    assert len(ops.Srsh_nnz) == nph and len(ops.Srs0_nnz) == npl
    Ef = ops.Vf
    for d in range(Nd):
        S = ops.Srsh_db[d]
        assert np.array_equal(S, -S.T)
        B = np.diag(ops.Brs[d])
        assert np.allclose(S @ np.ones(Nh), -np.concatenate([np.zeros(Nq), B]), atol=1e-13)
        Q = np.diag(rd.wq) @ rd.Drst[d]
        assert np.allclose(Q + Q.T, Ef.T @ ops.Brs[d] @ Ef, atol=1e-13)
        # the volume block of Srsh_db is Q - Q^T and the face block is E^T B
        assert np.allclose(S[:Nq, :Nq], Q - Q.T, atol=1e-14)
        assert np.allclose(S[:Nq, Nq:], Ef.T @ ops.Brs[d], atol=1e-14)
    # Vf interpolates degree-N polynomials to the face points exactly
    rf = np.array([-1.0, 1.0]) if Nd == 1 else None
    for p in range(N + 1):
        if Nd == 1:
            assert np.allclose(Ef @ rd.r ** p, rf ** p, atol=1e-14)
        else:
            for qd in range(N + 1):
                g = rd.r ** p * rd.s ** qd
                e = np.ones(N1D)
                fr = np.concatenate([-e, e, rd.r1D, rd.r1D])
                fs = np.concatenate([rd.r1D, rd.r1D, -e, e])
                assert np.allclose(Ef @ g, fr ** p * fs ** qd, atol=1e-13)
    assert np.allclose(ops.MinvVhT, np.diag(1 / ops.wq) @ np.hstack([np.eye(Nq), Ef.T]), rtol=1e-15, atol=0)
    assert np.allclose(ops.MinvVfT, np.diag(1 / ops.wq) @ Ef.T, rtol=1e-15, atol=0)
    assert np.array_equal(np.sort(np.unique(ops.Vf_low)), [0.0, 1.0]) and np.all(ops.Vf_low.sum(1) == 1.0)
    fmap = ops_mod.face_node_to_volume_node(N1D, Nd)
    assert np.array_equal(ops.fq2q - 1, fmap)
    for f, q in enumerate(fmap):
        assert (f + 1) in ops.q2fq[q]
