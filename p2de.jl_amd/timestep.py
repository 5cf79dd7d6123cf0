"""Host-side time-step estimate used to choose ``dt0`` for a problem.

``cfl_dt(fp, U)`` evaluates the low-order CFL time step of ``calculate_lambda_and_low_order_CFL!``
(low_order_graph_viscosity.jl:222-285) on the host with numpy:

    dt = min_{k,i} CFL * wq_i * J_ik / 2 / (sum_j lambda_ij + sum_{f in q2fq[i]} lambda_B_f),
    lambda_ij  = |S0_ij| max(ws(u_i, n_ij), ws(u_j, n_ji))            (rhs_utils.jl:105-128)
    lambda_B_f = |B_f| / 2 max(ws(u_f, n_f), ws(u_P, n_P))            (:283-285, :184-198)

with the reference's wave speed of the rotated 1D state (compressible_Navier_Stokes.jl:48-62).
It sets up a problem (``dt0 = 0.9 * cfl_dt / CFL``, so that the step cap binds at the first
step); the solvers compute their own time steps on the device.
"""
from __future__ import annotations

import numpy as np

from .dg import FlatProblem


def _wavespeed(gamma, U, n):
    """wavespeed_estimate of the states U (..., Nc) along unit normals n (..., 2) (2D) or any (1D)."""
    if U.shape[-1] == 3:
        m = U[..., 1]
        E = U[..., 2]
    else:
        m = n[..., 0] * U[..., 1] + n[..., 1] * U[..., 2]
        E = U[..., 3]
    rho = U[..., 0]
    p = (gamma - 1.0) * (E - 0.5 * m * m / rho)
    return np.abs(m / rho) + np.sqrt(gamma * p / rho)


def cfl_dt(fp: FlatProblem, U: np.ndarray) -> float:
    """Low-order CFL time step of the state U (shape (K, Nq, Nc)) without the CFL*dt0 / T cap."""
    K, Nq, Nc, Nfp, Nh, Nd = fp.K, fp.Nq, fp.Nc, fp.Nfp, fp.Nh, fp.Nd
    U = np.asarray(U, dtype=np.float64).reshape(K, Nq, Nc)
    J = fp.Jq.reshape(K, Nq)                                      # column-major (Nq, K) -> (K, Nq)
    G = fp.GJh.reshape(-1, K, Nh)[:, :, 0]                        # metric per element (constant)
    lam = np.zeros((K, Nq))
    # volume pairs (Srs0_nnz, i > j)
    pairs = fp.Srs0_nnz.reshape(-1, 2) - 1
    Sr0 = fp.Srs0[:Nq * Nq].reshape(Nq, Nq).T                     # column-major blocks
    Ss0 = fp.Srs0[Nq * Nq:].reshape(Nq, Nq).T if Nd == 2 else None
    for i, j in pairs:
        if Nd == 1:
            Sx = G[0] * Sr0[i, j]
            nrm = np.abs(Sx)
            n = np.zeros((K, 2))
            n[:, 0] = Sx / nrm
        else:
            Sx = G[0] * Sr0[i, j] + G[1] * Ss0[i, j]
            Sy = G[2] * Sr0[i, j] + G[3] * Ss0[i, j]
            nrm = np.hypot(Sx, Sy)
            n = np.stack([Sx / nrm, Sy / nrm], axis=-1)
        l = nrm * np.maximum(_wavespeed(fp.gamma, U[:, i], n), _wavespeed(fp.gamma, U[:, j], -n))
        lam[:, i] += l
        lam[:, j] += l
    # surface: lambda_B on each face node, added to its volume node
    Br = np.diag(fp.Brs[:Nfp * Nfp].reshape(Nfp, Nfp).T)
    Bs = np.diag(fp.Brs[Nfp * Nfp:].reshape(Nfp, Nfp).T) if Nd == 2 else np.zeros(Nfp)
    fq2q = fp.fq2q - 1
    mapP = fp.mapP.reshape(K, Nfp) - 1                            # (Nfp, K) column-major -> (K, Nfp)
    kP, iP = mapP // Nfp, mapP % Nfp
    UP = U[kP, fq2q[iP]]                                          # (K, Nfp, Nc)
    Uf = U[:, fq2q]
    if Nd == 1:
        Bx = G[0][:, None] * Br[None, :]
        nrm = np.abs(Bx)
        n = np.zeros((K, Nfp, 2))
        n[..., 0] = Bx / nrm
        BxP = G[0][kP] * Br[iP]
        nP = np.zeros((K, Nfp, 2))
        nP[..., 0] = BxP / np.abs(BxP)
    else:
        Bx = G[0][:, None] * Br[None, :] + G[1][:, None] * Bs[None, :]
        By = G[2][:, None] * Br[None, :] + G[3][:, None] * Bs[None, :]
        nrm = np.hypot(Bx, By)
        n = np.stack([Bx / nrm, By / nrm], axis=-1)
        BxP = G[0][kP] * Br[iP] + G[1][kP] * Bs[iP]
        ByP = G[2][kP] * Br[iP] + G[3][kP] * Bs[iP]
        nP = np.stack([BxP, ByP], axis=-1) / np.hypot(BxP, ByP)[..., None]
    lB = 0.5 * nrm * np.maximum(_wavespeed(fp.gamma, Uf, n), _wavespeed(fp.gamma, UP, nP))
    q2fq = fp.q2fq.reshape(2, Nq).T                               # (Nq, 2), 1-based, 0 = absent
    for s in range(2):
        has = q2fq[:, s] > 0
        f = q2fq[has, s] - 1
        lam[:, has] += lB[:, f]
    wq = fp.wq[None, :]
    return float(np.min(fp.CFL * 0.5 * wq * J / lam))
