"""Host-side mirror of the reference's set-up and state containers.

``initialize_DG(param, initial_condition, initial_boundary_conditions)`` returns
``(solver, state, state_param)`` like init.jl:44-60.  The containers keep the
reference's memory layout: a Julia ``Array{SVector{Nc,Float64},2}(Nq, K)`` is a numpy
array of shape ``(K, Nq, Nc)`` (component fastest, then node, then element), which is
byte-for-byte the layout the C-ABI reads (``include/p2de_hip.h``).

``flat_problem`` packs operators, geometry and boundary data into the contiguous,
column-major, 1-based arrays that ``p2de_create`` takes.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace
from typing import Callable, Optional

import numpy as np

from .operators import (BCData, MeshData, Operators, RefElem, SizeData, build_operators,
                        gauss_ref_elem, lobatto_ref_elem, uniform_mesh)
from .types import (GaussCollocation, LaxFriedrichsOnProjectedVal, NodewiseScaledExtrapolation, Param,
                    check_supported, limiter_codes)


@dataclass
class Discretization:
    sizes: SizeData
    geom: MeshData
    ops: Operators


@dataclass
class Solver:
    """``Solver`` (Solver.jl:220-225)."""
    param: Param
    rd: RefElem
    md: MeshData
    discrete_data: Discretization


@dataclass
class Preallocation:
    """The part of ``Preallocation`` (State.jl:1-26) the drop-in exposes on the host."""
    Uq: np.ndarray                       # (K, Nq, Nc)
    rhsU: np.ndarray                     # (K, Nq, Nc)
    rhsH: Optional[np.ndarray] = None
    rhsL: Optional[np.ndarray] = None
    L_local: Optional[np.ndarray] = None


@dataclass
class State:
    """``State`` (State.jl:258-261).  ``device`` holds the native context once created."""
    preallocation: Preallocation
    device: object = None


@dataclass
class StateParam:
    """``StateParam`` (StateParam.jl:9-11)."""
    bcdata: BCData


@dataclass
class TimeParam:
    """``TimeParam`` (TimeParam.jl:1-6) without the timer."""
    t: float
    dt: float
    nstage: int


def make_periodic(md: MeshData) -> MeshData:
    """StartUpDG ``make_periodic``: rebuild the face connectivity with wrap-around."""
    Kdims = md.Kdims
    periodic = True
    # rebuild connectivity only; coordinates and geometry are unchanged
    Nd = md.Nd
    Nfp = md.mapP.shape[0]
    N1D = Nfp // 4 if Nd == 2 else None
    if Nd == 1:
        K = Kdims[0]
        mapP = md.mapP.copy() - 1
        mapP[0, 0] = 1 + 2 * (K - 1)
        mapP[1, K - 1] = 0
        return replace(md, mapP=mapP + 1, boundary=np.zeros_like(md.boundary))
    Kx, Ky = Kdims
    from .operators import quad_connectivity
    mapP, _ = quad_connectivity(Kx, Ky, N1D, True, True)
    return replace(md, mapP=mapP + 1, boundary=np.zeros_like(md.boundary))


def initialize_DG(param: Param, initial_condition: Callable, initial_boundary_conditions: Callable):
    """``initialize_DG`` (init.jl:44-60).

    ``initial_condition(param, x[, y])`` is called with node-coordinate arrays of shape
    (Nq, K) and returns conservative variables of shape (Nc, Nq, K);
    ``initial_boundary_conditions(param, md)`` returns a ``BCData``.
    """
    check_supported(param)
    Nd = param.equation.Nd
    if isinstance(param.approximation_basis, GaussCollocation):    # init.jl:66-72, 82-95
        rd = gauss_ref_elem(param.N, Nd)
        ops, sizes = build_operators(rd, param.global_constants.ZEROTOL, quad="gauss")
    else:                                                          # init.jl:74-80, 97-103
        rd = lobatto_ref_elem(param.N, Nd)
        ops, sizes = build_operators(rd, param.global_constants.ZEROTOL)
    md = uniform_mesh(rd, param.K, param.xL, param.xR, periodic=False)
    sizes.K = param.num_elements
    bcdata = initial_boundary_conditions(param, md)
    if isinstance(bcdata, tuple):      # (BCData, mesh with periodic connectivity)
        bcdata, md = bcdata
    discrete = Discretization(sizes=sizes, geom=md, ops=ops)
    solver = Solver(param=param, rd=rd, md=md, discrete_data=discrete)
    # init_U! (init.jl:342-361)
    if Nd == 1:
        U = np.asarray(initial_condition(param, md.xq), dtype=np.float64)
    else:
        U = np.asarray(initial_condition(param, md.xq, md.yq), dtype=np.float64)
    Nc = param.equation.Nc
    if U.shape != (Nc, sizes.Nq, sizes.K):
        raise ValueError(f"initial_condition must return shape {(Nc, sizes.Nq, sizes.K)}, got {U.shape}")
    Uq = np.ascontiguousarray(np.transpose(U, (2, 1, 0)))
    state = State(preallocation=Preallocation(Uq=Uq, rhsU=np.zeros_like(Uq)))
    return solver, state, StateParam(bcdata=bcdata)


def primitive_to_conservative(gamma: float, prim):
    """``primitive_to_conservative`` (compressible_Navier_Stokes.jl:1-16); prim is
    (rho, u, p) or (rho, u, v, p) of equal-shape arrays; returns a stacked array."""
    if len(prim) == 3:
        rho, u, p = prim
        rhou = rho * u
        E = p / (gamma - 1) + 0.5 * rho * (u * u)
        return np.stack([rho, rhou, E])
    rho, u, v, p = prim
    rhou = rho * u
    rhov = rho * v
    E = p / (gamma - 1) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rhou, rhov, E])


# ---------------------------------------------------------------------------------
# flat problem for the C-ABI
# ---------------------------------------------------------------------------------
def _colmajor(A: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(A, dtype=np.float64).T).ravel()


@dataclass
class FlatProblem:
    """Contiguous arrays in the layout of ``p2de_operators``/``p2de_geometry``/``p2de_bcdata``."""
    Nd: int
    Nc: int
    N1D: int
    Nq: int
    Nfp: int
    Nh: int
    K: int
    Kx: int
    Ky: int
    gamma: float
    zeta: float
    eta: float
    POSTOL: float
    ZEROTOL: float
    CFL: float
    dt0: float
    t0: float
    T: float
    Srsh_db: np.ndarray
    Srs0: np.ndarray
    Srsh_nnz: np.ndarray
    Srs0_nnz: np.ndarray
    Brs: np.ndarray
    MinvVhT: np.ndarray
    MinvVfT: np.ndarray
    Vf: np.ndarray
    Vf_low: np.ndarray
    wq: np.ndarray
    fq2q: np.ndarray
    q2fq: np.ndarray
    Jq: np.ndarray
    GJh: np.ndarray
    mapP: np.ndarray
    mapI: np.ndarray
    Ival: np.ndarray
    mapO: np.ndarray
    upper_bound_mode: int = 0
    basis: int = 0                  # 0 LobattoCollocation, 1 GaussCollocation
    entropyproj_limiter: int = 0    # 0 NoEntropyProjectionLimiter, 1 NodewiseScaledExtrapolation
    rhs_limiter: int = 0            # 0 SubcellLimiter, 1 ZhangShuLimiter
    bound: int = 0                  # 0 PositivityBound, 1 ...MinEntropyBound, 2 ...RelaxedMinEntropyBound,
                                    # 3 ...CellEntropyBound, 4 ...RelaxedCellEntropyBound
    shockcapture: int = 0           # 0 NoShockCapture, 1 HennemannShockCapture(shock_a, shock_c)
    shock_a: float = 0.5
    shock_c: float = 1.8
    cell_beta: float = 0.5          # PositivityAndRelaxedCellEntropyBound(beta)
    VDM_inv: Optional[np.ndarray] = None
    low_order_surface_flux: int = 0   # 0 LaxFriedrichsOnNodalVal, 1 LaxFriedrichsOnProjectedVal


def flat_problem(solver: Solver, state_param: StateParam, upper_bound_mode: int = 0) -> FlatProblem:
    p = solver.param
    ops = solver.discrete_data.ops
    md = solver.md
    sz = solver.discrete_data.sizes
    Nd = sz.Nd
    q2fq = np.zeros((sz.Nq, 2), dtype=np.int64)
    for i, fl in enumerate(ops.q2fq):
        for m, f in enumerate(fl[:2]):
            q2fq[i, m] = f
    bc = state_param.bcdata
    Kx, Ky = (sz.K, 1) if Nd == 1 else (int(p.K[0]), int(p.K[1]))
    Ival = np.asarray(bc.Ival, dtype=np.float64).reshape(-1, sz.Nc) if len(bc.mapI) else np.zeros((0, sz.Nc))
    tp = p.timestepping_param
    lc = limiter_codes(p.rhs_limiter)
    return FlatProblem(
        Nd=Nd, Nc=sz.Nc, N1D=sz.N1D, Nq=sz.Nq, Nfp=sz.Nfp, Nh=sz.Nh, K=sz.K, Kx=Kx, Ky=Ky,
        gamma=float(p.equation.gamma), zeta=float(p.limiting_param.zeta), eta=float(p.limiting_param.eta),
        POSTOL=float(p.global_constants.POSTOL), ZEROTOL=float(p.global_constants.ZEROTOL),
        CFL=float(tp.CFL), dt0=float(tp.dt0), t0=float(tp.t0), T=float(tp.T),
        Srsh_db=np.concatenate([_colmajor(S) for S in ops.Srsh_db]),
        Srs0=np.concatenate([_colmajor(S) for S in ops.Srs0]),
        Srsh_nnz=np.ascontiguousarray(np.array(ops.Srsh_nnz, dtype=np.int64).reshape(-1, 2)).ravel(),
        Srs0_nnz=np.ascontiguousarray(np.array(ops.Srs0_nnz, dtype=np.int64).reshape(-1, 2)).ravel(),
        Brs=np.concatenate([_colmajor(B) for B in ops.Brs]),
        MinvVhT=_colmajor(ops.MinvVhT), MinvVfT=_colmajor(ops.MinvVfT),
        Vf=_colmajor(ops.Vf), Vf_low=_colmajor(ops.Vf_low), wq=np.ascontiguousarray(ops.wq, dtype=np.float64),
        fq2q=np.ascontiguousarray(ops.fq2q, dtype=np.int64),
        q2fq=np.ascontiguousarray(q2fq.T).ravel(),
        Jq=_colmajor(md.J), GJh=np.concatenate([_colmajor(G) for G in md.GJh]),
        mapP=np.ascontiguousarray(np.asarray(bc.mapP, dtype=np.int64).T).ravel(),
        mapI=np.ascontiguousarray(np.asarray(bc.mapI, dtype=np.int64)),
        Ival=np.ascontiguousarray(Ival).ravel(),
        mapO=np.ascontiguousarray(np.asarray(bc.mapO, dtype=np.int64)),
        upper_bound_mode=int(upper_bound_mode),
        basis=1 if isinstance(p.approximation_basis, GaussCollocation) else 0,
        entropyproj_limiter=1 if isinstance(p.entropyproj_limiter, NodewiseScaledExtrapolation) else 0,
        rhs_limiter=lc[0], bound=lc[1], shockcapture=lc[2], shock_a=lc[3], shock_c=lc[4], cell_beta=lc[5],
        VDM_inv=_colmajor(ops.VDM_inv),
        low_order_surface_flux=1 if isinstance(p.rhs.low_order_surface_flux, LaxFriedrichsOnProjectedVal) else 0)
