"""Reference element, hybridized SBP operators, uniform mesh and face connectivity.

Host-side set-up, computed once in FP64 with numpy.  It restates what the reference
obtains from StartUpDG/NodesAndModes in ``initialize_operators``
(src/dg/init.jl:133-228) for the Lobatto (SBP) collocation basis and for the Gauss
collocation basis (``construct_gauss_reference_data``, init.jl:105-131):

* 1D Legendre-Gauss-Lobatto nodes/weights and the Lagrange differentiation matrix;
* 2D tensor-product nodes, r index fastest (init.jl:87, limiter_utils.jl:97-120);
* faces ordered [r=-1, r=+1, s=-1, s=+1], N+1 nodes each, ordered by the other
  coordinate ascending (subcell.jl:173-206, init.jl:330-333);
* ``Qrsh``, ``Srsh_db``, ``Srs0``, the nonzero pair lists, ``fq2q``/``q2fq``
  (init.jl:147-220, 276-322) with the reference's ``droptol`` and 1.0-snapping;
* uniform meshes with elements x-fastest (postprocess.jl:169-171), affine geometric
  factors, and ``mapP``/``mapI``/``mapO`` in the reference's 1-based linear indexing
  into (Nfp, K) (StateParam.jl:1-7).

StartUpDG computes the same matrices through Vandermonde solves, so its entries can
differ from these in the last bits; nothing here can pin that (no Julia on this host).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np


# ---------------------------------------------------------------------------------
# 1D nodes and operators
# ---------------------------------------------------------------------------------
def lgl_nodes_weights(N: int) -> Tuple[np.ndarray, np.ndarray]:
    """Legendre-Gauss-Lobatto nodes (ascending) and weights for degree N (N+1 points).

    Newton iteration on (1-x^2) P_N'(x) from Chebyshev-Gauss-Lobatto guesses, weights
    w_i = 2 / (N (N+1) P_N(x_i)^2).  Nodes are symmetrised so x_i = -x_{N-i} exactly.
    """
    if N < 1:
        raise ValueError("LGL needs N >= 1")
    n = N + 1
    x = np.cos(np.pi * np.arange(n) / N)
    P = np.zeros((n, n))
    xold = np.full(n, 2.0)
    for _ in range(100):
        if np.max(np.abs(x - xold)) <= 1e-16:
            break
        xold = x.copy()
        P[:, 0] = 1.0
        P[:, 1] = x
        for k in range(2, n):
            P[:, k] = ((2 * k - 1) * x * P[:, k - 1] - (k - 1) * P[:, k - 2]) / k
        x = xold - (x * P[:, N] - P[:, N - 1]) / (n * P[:, N])
    x = np.sort(x)
    x = 0.5 * (x - x[::-1])           # exact antisymmetry
    x[0], x[-1] = -1.0, 1.0
    if n % 2 == 1:
        x[N // 2] = 0.0
    # Legendre P_N at the final nodes for the weights
    p0, p1 = np.ones(n), x.copy()
    for k in range(2, n):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    PN = p1 if N >= 1 else p0
    w = 2.0 / (N * (N + 1) * PN * PN)
    w = 0.5 * (w + w[::-1])
    return x, w


def gauss_nodes_weights(N: int) -> Tuple[np.ndarray, np.ndarray]:
    """Legendre-Gauss nodes (ascending) and weights, N+1 points: ``gauss_quad(0, 0, N)``.

    numpy's ``leggauss`` (Golub-Welsch eigenvalues of the Jacobi matrix, one Newton step on
    P_{N+1}); symmetrised so x_i = -x_{N-i} and w_i = w_{N-i} exactly.
    """
    x, w = np.polynomial.legendre.leggauss(N + 1)
    x = np.asarray(x, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    x = 0.5 * (x - x[::-1])
    if (N + 1) % 2 == 1:
        x[N // 2] = 0.0
    w = 0.5 * (w + w[::-1])
    return x, w


def lagrange_interp_row(x: np.ndarray, xi: float) -> np.ndarray:
    """Row l_a(xi), a = 0..n-1, of the Lagrange basis on nodes x (barycentric form)."""
    n = len(x)
    lam = np.array([1.0 / np.prod([x[j] - x[k] for k in range(n) if k != j]) for j in range(n)])
    d = xi - x
    if np.any(d == 0.0):
        return (d == 0.0).astype(np.float64)
    t = lam / d
    return t / np.sum(t)


def lagrange_diff_matrix(x: np.ndarray) -> np.ndarray:
    """D_ij = l_j'(x_i) by barycentric weights; diagonal by the negative-sum rule."""
    n = len(x)
    lam = np.array([1.0 / np.prod([x[j] - x[k] for k in range(n) if k != j]) for j in range(n)])
    D = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i != j:
                D[i, j] = (lam[j] / lam[i]) / (x[i] - x[j])
        D[i, i] = -np.sum(D[i, np.arange(n) != i])
    return D


def legendre_orthonormal(x: np.ndarray, n: int) -> np.ndarray:
    """P_n(x) scaled to unit L2 norm on [-1, 1] (NodesAndModes ``jacobiP(x, 0, 0, n)``)."""
    p0 = np.ones_like(x)
    if n == 0:
        return p0 * np.sqrt(0.5)
    p1 = x.copy()
    for k in range(2, n + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    return p1 * np.sqrt((2 * n + 1) / 2.0)


def vandermonde(rd) -> np.ndarray:
    """``VDM`` of the element (init.jl:116,141): the orthonormal Legendre basis at the volume
    nodes; in 2D mode (i, j) = P_i(r) P_j(s) in column j + (N+1) i (NodesAndModes' order)."""
    N = rd.N
    if rd.Nd == 1:
        return np.stack([legendre_orthonormal(rd.r, i) for i in range(N + 1)], axis=1)
    cols = []
    for i in range(N + 1):
        for j in range(N + 1):
            cols.append(legendre_orthonormal(rd.r, i) * legendre_orthonormal(rd.s, j))
    return np.stack(cols, axis=1)


def droptol(A: np.ndarray, tol: float) -> np.ndarray:
    """SparseArrays ``droptol!``: keep entries with ``abs(x) > tol``."""
    B = np.array(A, dtype=np.float64, copy=True)
    B[np.abs(B) <= tol] = 0.0
    return B


# ---------------------------------------------------------------------------------
# reference element
# ---------------------------------------------------------------------------------
@dataclass
class RefElem:
    """The subset of StartUpDG's ``RefElemData`` that init.jl:141 reads."""
    Nd: int
    N: int
    r: np.ndarray                 # (Nq,) volume node r coordinates
    s: Optional[np.ndarray]       # (Nq,) s coordinates (2D)
    wq: np.ndarray                # (Nq,)
    wf: np.ndarray                # (Nfp,)
    nrJ: np.ndarray               # (Nfp,)
    nsJ: Optional[np.ndarray]     # (Nfp,)
    Drst: Tuple[np.ndarray, ...]  # Dr (, Ds)
    Vf: np.ndarray                # (Nfp, Nq) interpolation to face nodes
    r1D: np.ndarray
    w1D: np.ndarray


def face_node_to_volume_node(N1D: int, Nd: int) -> np.ndarray:
    """0-based volume node of each 0-based face node (LGL: faces pass through nodes)."""
    if Nd == 1:
        return np.array([0, N1D - 1])
    f = []
    for m in range(N1D):
        f.append(m * N1D)                     # r = -1 (left), s index m
    for m in range(N1D):
        f.append(m * N1D + N1D - 1)           # r = +1 (right)
    for m in range(N1D):
        f.append(m)                           # s = -1 (bottom), r index m
    for m in range(N1D):
        f.append((N1D - 1) * N1D + m)         # s = +1 (top)
    return np.array(f)


def lobatto_ref_elem(N: int, Nd: int) -> RefElem:
    """``RefElemData(Line(), N, quad_rule_vol=gauss_lobatto_quad)`` (init.jl:74-80) and
    ``RefElemData(Quad(), SBP(), N)`` (init.jl:97-103): collocated LGL nodes."""
    r1D, w1D = lgl_nodes_weights(N)
    D1D = lagrange_diff_matrix(r1D)
    N1D = N + 1
    fmap = face_node_to_volume_node(N1D, Nd)
    if Nd == 1:
        Nq = N1D
        Vf = np.zeros((2, Nq))
        Vf[0, 0] = 1.0
        Vf[1, -1] = 1.0
        return RefElem(1, N, r1D.copy(), None, w1D.copy(), np.ones(2), np.array([-1.0, 1.0]), None,
                       (D1D,), Vf, r1D, w1D)
    Nq = N1D * N1D
    idx = np.arange(Nq)
    i_r, j_s = idx % N1D, idx // N1D
    r, s = r1D[i_r], r1D[j_s]
    wq = w1D[i_r] * w1D[j_s]                  # wq = wr .* ws (init.jl:88-89)
    I = np.eye(N1D)
    Dr = np.kron(I, D1D)                      # acts on the r index (fastest)
    Ds = np.kron(D1D, I)
    Nfp = 4 * N1D
    Vf = np.zeros((Nfp, Nq))
    Vf[np.arange(Nfp), fmap] = 1.0
    e = np.ones(N1D)
    z = np.zeros(N1D)
    nrJ = np.concatenate([-e, e, z, z])
    nsJ = np.concatenate([z, z, -e, e])
    wf = np.concatenate([w1D, w1D, w1D, w1D])
    return RefElem(2, N, r, s, wq, wf, nrJ, nsJ, (Dr, Ds), Vf, r1D, w1D)


def gauss_ref_elem(N: int, Nd: int) -> RefElem:
    """``construct_gauss_reference_data`` (init.jl:105-131) of the Gauss-collocation element
    (init.jl:66-72 in 1D, :82-95 in 2D): Lagrange basis on the N+1 Legendre-Gauss nodes per
    direction (r fastest), Vq = Pq = I, M = diag(wq), Drst = the Gauss-node differentiation
    matrices, Vf = interpolation along each line to the face points, which are the Gauss
    points of the faces (``quad_rule_face=(r1D, w1D)``).

    StartUpDG forms these through the LGL nodal basis (Vf = rd.Vf * inv(Vq) etc.); the
    matrices are the same in exact arithmetic and are computed here directly.
    """
    r1D, w1D = gauss_nodes_weights(N)
    D1D = lagrange_diff_matrix(r1D)
    lm = lagrange_interp_row(r1D, -1.0)
    lp = lagrange_interp_row(r1D, 1.0)
    N1D = N + 1
    if Nd == 1:
        Vf = np.vstack([lm, lp])
        return RefElem(1, N, r1D.copy(), None, w1D.copy(), np.ones(2), np.array([-1.0, 1.0]), None,
                       (D1D,), Vf, r1D, w1D)
    Nq = N1D * N1D
    idx = np.arange(Nq)
    i_r, j_s = idx % N1D, idx // N1D
    r, s = r1D[i_r], r1D[j_s]
    wq = w1D[i_r] * w1D[j_s]
    I = np.eye(N1D)
    Dr = np.kron(I, D1D)
    Ds = np.kron(D1D, I)
    Nfp = 4 * N1D
    Vf = np.zeros((Nfp, Nq))
    a = np.arange(N1D)
    for m in range(N1D):
        Vf[m, m * N1D + a] = lm                   # r = -1: the r-line at s index m
        Vf[N1D + m, m * N1D + a] = lp             # r = +1
        Vf[2 * N1D + m, a * N1D + m] = lm         # s = -1: the s-line at r index m
        Vf[3 * N1D + m, a * N1D + m] = lp         # s = +1
    e = np.ones(N1D)
    z = np.zeros(N1D)
    nrJ = np.concatenate([-e, e, z, z])
    nsJ = np.concatenate([z, z, -e, e])
    wf = np.concatenate([w1D, w1D, w1D, w1D])
    return RefElem(2, N, r, s, wq, wf, nrJ, nsJ, (Dr, Ds), Vf, r1D, w1D)


def gauss_low_order_extrapolation(N: int, Nd: int) -> np.ndarray:
    """``Vf_low`` of the Gauss element: 1.0 at the line-end node of each face point
    (init.jl:324-336 in 2D; ``construct_low_order_operators_1D`` init.jl:288-290 in 1D)."""
    N1D = N + 1
    fmap = face_node_to_volume_node(N1D, Nd)
    Nq = N1D ** Nd
    V = np.zeros((len(fmap), Nq))
    V[np.arange(len(fmap)), fmap] = 1.0
    return V


# ---------------------------------------------------------------------------------
# operators (init.jl:133-228, 276-322)
# ---------------------------------------------------------------------------------
@dataclass
class Operators:
    """Mirror of ``Operators{DIM}`` (Solver.jl:192-212); arrays are FP64, 0-based
    numpy, and ``*_nnz``/``fq2q``/``q2fq`` hold 1-based indices as in the reference."""
    Srsh_db: Tuple[np.ndarray, ...]   # Nd x (Nh, Nh)
    Srs0: Tuple[np.ndarray, ...]      # Nd x (Nq, Nq)
    Srsh_nnz: List[Tuple[int, int]]
    Srs0_nnz: List[Tuple[int, int]]
    Brs: Tuple[np.ndarray, ...]       # Nd x (Nfp, Nfp) diagonal
    Vh: np.ndarray
    MinvVhT: np.ndarray               # (Nq, Nh)
    MinvVfT: np.ndarray               # (Nq, Nfp)
    Vq: np.ndarray
    Vf: np.ndarray                    # (Nfp, Nq)
    Vf_low: np.ndarray                # (Nfp, Nq)
    Pq: np.ndarray
    wq: np.ndarray
    q2fq: List[List[int]]
    fq2q: np.ndarray
    VDM_inv: Optional[np.ndarray] = None   # (Np, Np): inv(VDM) (Solver.jl:200, init.jl:224)


@dataclass
class SizeData:
    """``SizeData`` (Solver.jl:173-183)."""
    K: int
    N1D: int
    Nd: int
    Nc: int
    Np: int
    Nq: int
    Nfp: int
    Nh: int
    Ns: int = 3


def low_order_Q01D(N: int, ZEROTOL: float) -> np.ndarray:
    """Tridiagonal low-order operator of ``construct_low_order_operators_1D`` (init.jl:276-295)."""
    Nq = N + 1
    Q = np.zeros((Nq, Nq))
    for i in range(Nq - 1):
        Q[i + 1, i] = -0.5      # dl
        Q[i, i + 1] = 0.5       # du
    Q[0, 0] = -0.5
    Q[-1, -1] = 0.5
    return droptol(Q, ZEROTOL)


def build_operators(rd: RefElem, ZEROTOL: float, quad: str = "lobatto") -> Tuple[Operators, SizeData]:
    """``initialize_operators`` minus the mesh (init.jl:141-220); ``quad`` is the
    quadrature the low-order extrapolation dispatches on (init.jl:302-340)."""
    Nd = rd.Nd
    Nq = len(rd.wq)
    Nfp = rd.Vf.shape[0]
    Nh = Nq + Nfp
    M = np.diag(rd.wq)
    Pq = np.eye(Nq)
    Vq = np.eye(Nq)
    Vf = rd.Vf.copy()
    Qrs = tuple(droptol(Pq.T @ M @ D @ Pq, ZEROTOL) for D in rd.Drst)
    Ef = Vf @ Pq
    nrst = (rd.nrJ,) if Nd == 1 else (rd.nrJ, rd.nsJ)
    Brs = tuple(droptol(np.diag(rd.wf * n), ZEROTOL) for n in nrst)
    Qrsh = []
    for Q, B in zip(Qrs, Brs):
        top = np.hstack([Q - Q.T, Ef.T @ B])
        bot = np.hstack([-B @ Ef, B])
        Qrsh.append(droptol(0.5 * np.vstack([top, bot]), ZEROTOL))
    Srsh = [0.5 * (A - A.T) for A in Qrsh]
    Srsh_db = tuple(2.0 * A for A in Srsh)

    N1D = rd.N + 1
    Nc = Nd + 2
    sizes = SizeData(K=0, N1D=N1D, Nd=Nd, Nc=Nc, Np=Nq, Nq=Nq, Nfp=Nfp, Nh=Nh)

    Vh = droptol(np.vstack([Vq, Vf]), ZEROTOL)
    MinvVhT = np.linalg.solve(M, Vh.T)
    MinvVfT = np.linalg.solve(M, Vf.T)

    # low-order operators (init.jl:297-322)
    Q01D = low_order_Q01D(rd.N, ZEROTOL)
    if Nd == 1:
        Srs0 = (droptol(0.5 * (Q01D - Q01D.T), ZEROTOL),)
    else:
        M1D = np.diag(rd.w1D)
        Qr0 = droptol(np.kron(M1D, Q01D), ZEROTOL)
        Qs0 = droptol(np.kron(Q01D, M1D), ZEROTOL)
        Srs0 = (droptol(0.5 * (Qr0 - Qr0.T), ZEROTOL), droptol(0.5 * (Qs0 - Qs0.T), ZEROTOL))
    if quad == "gauss":
        Vf_low = gauss_low_order_extrapolation(rd.N, Nd)
    else:
        Vf_low = droptol(Vf, ZEROTOL)
    Vf_low[np.abs(Vf_low - 1.0) < ZEROTOL] = 1.0          # init.jl:181-185

    Srsh_nnz = [(i + 1, j + 1) for j in range(Nh) for i in range(j + 1, Nh)
                if sum(abs(S[i, j]) for S in Srsh_db) != 0]
    Srs0_nnz = [(i + 1, j + 1) for j in range(Nq) for i in range(j + 1, Nq)
                if sum(abs(S[i, j]) for S in Srs0) != 0]
    fq2q = np.array([int(np.flatnonzero(Vf_low[i, :] == 1.0)[0]) + 1 for i in range(Nfp)], dtype=np.int64)
    q2fq = [[int(f) + 1 for f in np.flatnonzero(Vf_low[:, i] == 1.0)] for i in range(Nq)]
    ops = Operators(Srsh_db=Srsh_db, Srs0=Srs0, Srsh_nnz=Srsh_nnz, Srs0_nnz=Srs0_nnz, Brs=Brs,
                    Vh=Vh, MinvVhT=MinvVhT, MinvVfT=MinvVfT, Vq=Vq, Vf=Vf, Vf_low=Vf_low, Pq=Pq,
                    wq=rd.wq.copy(), q2fq=q2fq, fq2q=fq2q, VDM_inv=np.linalg.inv(vandermonde(rd)))
    return ops, sizes


# ---------------------------------------------------------------------------------
# mesh, geometry, connectivity (init.jl:230-274; StartUpDG uniform_mesh/make_periodic)
# ---------------------------------------------------------------------------------
@dataclass
class MeshData:
    """Node coordinates, geometric factors and face connectivity of a uniform mesh."""
    Nd: int
    Kdims: Tuple[int, ...]             # (K,) or (Kx, Ky)
    xq: np.ndarray                     # (Nq, K)
    yq: Optional[np.ndarray]           # (Nq, K)
    J: np.ndarray                      # (Nq, K)
    GJh: Tuple[np.ndarray, ...]        # rxJh (, sxJh, ryJh, syJh), each (Nh, K)
    mapP: np.ndarray                   # (Nfp, K) int64, 1-based linear index into (Nfp, K)
    boundary: np.ndarray               # (Nfp, K) bool: face node on the domain boundary
    h: Tuple[float, ...]               # element sizes


def _vertices(xL: float, xR: float, K: int) -> np.ndarray:
    v = -1.0 + 2.0 * np.arange(K + 1) / K
    v[0], v[-1] = -1.0, 1.0
    return (v + 1.0) / 2.0 * (xR - xL) + xL


def uniform_mesh(rd: RefElem, Kdims, xL, xR, periodic) -> MeshData:
    """Uniform line/quad mesh.  ``periodic`` is a bool (1D) or a pair (2D); a
    non-periodic boundary face node maps to itself, as StartUpDG's ``mapP`` does."""
    N1D = rd.N + 1
    if rd.Nd == 1:
        K = int(Kdims)
        VX = _vertices(float(xL), float(xR), K)
        h = (float(xR) - float(xL)) / K
        r = rd.r
        xq = 0.5 * (1.0 - r)[:, None] * VX[None, :-1] + 0.5 * (1.0 + r)[:, None] * VX[None, 1:]
        Nq = N1D
        Nh = Nq + 2
        J = np.full((Nq, K), h / 2.0)
        rxJh = np.ones((Nh, K))
        mapP = np.zeros((2, K), dtype=np.int64)
        bnd = np.zeros((2, K), dtype=bool)
        for k in range(K):
            # face 0 (left) <-> face 1 (right) of element k-1
            kl, kr = k - 1, k + 1
            if kl < 0:
                kl = K - 1 if periodic else None
            if kr >= K:
                kr = 0 if periodic else None
            mapP[0, k] = (1 + 2 * kl) if kl is not None else (0 + 2 * k)
            mapP[1, k] = (0 + 2 * kr) if kr is not None else (1 + 2 * k)
            bnd[0, k] = kl is None
            bnd[1, k] = kr is None
        return MeshData(1, (K,), xq, None, J, (rxJh,), mapP + 1, bnd, (h,))

    Kx, Ky = int(Kdims[0]), int(Kdims[1])
    K = Kx * Ky
    px, py = (periodic, periodic) if isinstance(periodic, bool) else periodic
    VX = _vertices(float(xL[0]), float(xR[0]), Kx)
    VY = _vertices(float(xL[1]), float(xR[1]), Ky)
    hx = (float(xR[0]) - float(xL[0])) / Kx
    hy = (float(xR[1]) - float(xL[1])) / Ky
    ex = np.arange(K) % Kx
    ey = np.arange(K) // Kx
    r, s = rd.r, rd.s
    xq = 0.5 * (1.0 - r)[:, None] * VX[None, ex] + 0.5 * (1.0 + r)[:, None] * VX[None, ex + 1]
    yq = 0.5 * (1.0 - s)[:, None] * VY[None, ey] + 0.5 * (1.0 + s)[:, None] * VY[None, ey + 1]
    Nq = N1D * N1D
    Nfp = 4 * N1D
    Nh = Nq + Nfp
    # affine geometric factors: x_r = hx/2, y_s = hy/2 -> rxJ = y_s, syJ = x_r, J = x_r y_s
    xr, ys = hx / 2.0, hy / 2.0
    J = np.full((Nq, K), xr * ys)
    rxJh = np.full((Nh, K), ys)
    sxJh = np.zeros((Nh, K))
    ryJh = np.zeros((Nh, K))
    syJh = np.full((Nh, K), xr)
    mapP, bnd = quad_connectivity(Kx, Ky, N1D, px, py)
    return MeshData(2, (Kx, Ky), xq, yq, J, (rxJh, sxJh, ryJh, syJh), mapP + 1, bnd, (hx, hy))


def quad_connectivity(Kx: int, Ky: int, N1D: int, px: bool, py: bool):
    """0-based mapP (Nfp, K) of a Kx x Ky quad grid (elements x-fastest) and the boundary
    mask.  Face f = face*N1D + m: faces left/right/bottom/top; node m of a face matches node
    m of the opposite face of the neighbour; a boundary face node maps to itself."""
    K = Kx * Ky
    Nfp = 4 * N1D
    k = np.arange(K)
    ex, ey = k % Kx, k // Kx
    m = np.arange(N1D)
    mapP = np.empty((Nfp, K), dtype=np.int64)
    bnd = np.zeros((Nfp, K), dtype=bool)
    specs = [  # (face, opposite face, neighbour element, exists)
        (0, 1, (ex - 1) % Kx + ey * Kx, (ex > 0) | px),
        (1, 0, (ex + 1) % Kx + ey * Kx, (ex < Kx - 1) | px),
        (2, 3, ex + ((ey - 1) % Ky) * Kx, (ey > 0) | py),
        (3, 2, ex + ((ey + 1) % Ky) * Kx, (ey < Ky - 1) | py),
    ]
    for face, opp, nbr, ok in specs:
        f = face * N1D + m[:, None]
        target = opp * N1D + m[:, None] + Nfp * nbr[None, :]
        self_ = f + Nfp * k[None, :]
        okb = np.broadcast_to(np.asarray(ok)[None, :], (N1D, K))
        mapP[face * N1D:(face + 1) * N1D, :] = np.where(okb, target, self_)
        bnd[face * N1D:(face + 1) * N1D, :] = ~okb
    return mapP, bnd


@dataclass
class BCData:
    """``BCData`` (StateParam.jl:1-7): 1-based linear face-node indices."""
    mapP: np.ndarray
    mapI: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int64))
    mapO: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int64))
    Ival: np.ndarray = field(default_factory=lambda: np.zeros((0, 4)))
