"""The benchmark and test problems, written as the reference's driver scripts write them:
a ``Param``, an ``initial_condition(param, x[, y])`` and an
``initial_boundary_conditions(param, md)`` (test/test_smoke.jl:22-36,68-104,117-129).

Problems (γ = 1.4, ζ = 0.1, POSTOL = 1e-14, ZEROTOL = 5e-16 unless stated):

* ``sod_1d``          -- Sod tube on [0,1], K=100, N=3, CFL 0.5, T 0.2, transmissive ends.
* ``vortex_2d``       -- isentropic vortex on periodic [-5,5]^2, 64x64, N=3, strength 5,
                         free stream (1,1,1,1), T 1.0.
* ``riemann4_2d``     -- four-state Riemann problem on [0,1]^2, 256x256, N=3, T 0.25,
                         transmissive boundary.
* ``kelvin_helmholtz``-- periodic [-1,1]^2, 512x512, N=4 (examples/2D/kelvin-helmholtz.jl:9-18).
* ``kelvin_helmholtz_gauss`` -- the reference's KH example as written: 80x80, N=3, Gauss
                         collocation with NodewiseScaledExtrapolation, dt0 1e-3.
* ``astro_jet``       -- the reference's Mach-2000-class jet (examples/2D/astro-jet.jl): inflow on
                         the left, outflow on the right, gamma 5/3, Gauss + relaxed min-entropy bound.
* ``smoke_vortex``    -- the reference's own smoke problem (test/test_smoke.jl:6-20,117-129).
* ``periodic_wave``   -- one instance of the batched workload: periodic [0,1]^2, density
                         1 + 0.5 sin(2 pi kx x) sin(2 pi ky y), constant velocity, p = 1, with an
                         optional near-vacuum Gaussian dip (``batched_instances`` draws 48).

Every problem takes ``basis``: "lobatto" (default), "gauss" (GaussCollocation +
NodewiseScaledExtrapolation) or "gauss_noep" (GaussCollocation, no entropy-projection limiter).

``dt0`` for the first four is ``0.9 * dt_pos(U0) / CFL`` where ``dt_pos`` is the
low-order positivity time step of the initial state, recorded once (so the CFL cap
binds at the first step and the limiter's dt equals the update dt there).
"""
from __future__ import annotations

import math
from dataclasses import replace
from typing import Callable, Dict, Tuple

import numpy as np

from .dg import initialize_DG, make_periodic, primitive_to_conservative
from .operators import BCData
from .types import (CompressibleEulerIdealGas, Dim1, Dim2, ESLimitedLowOrderPos, GaussCollocation,
                    GlobalConstant, HennemannShockCapture, LaxFriedrichsOnNodalVal, LaxFriedrichsOnProjectedVal,
                    LimitingParameter, LobattoCollocation,
                    NodewiseScaledExtrapolation, NoEntropyProjectionLimiter, Param,
                    PositivityAndCellEntropyBound, PositivityAndMinEntropyBound,
                    PositivityAndRelaxedCellEntropyBound, PositivityAndRelaxedMinEntropyBound, PositivityBound,
                    SubcellLimiter, TimesteppingParameter, TVDAndCellEntropyBound, TVDAndMinEntropyBound,
                    TVDAndRelaxedCellEntropyBound, TVDAndRelaxedMinEntropyBound, TVDBound, ZhangShuLimiter)

# approximation basis and entropy-projection limiter by name
BASES = {
    "lobatto": (LobattoCollocation, NoEntropyProjectionLimiter),
    "gauss": (GaussCollocation, NodewiseScaledExtrapolation),        # examples/2D/kelvin-helmholtz.jl:52-53
    "gauss_noep": (GaussCollocation, NoEntropyProjectionLimiter),
}

# rhs limiters by name: the limiter set-ups of test/test_smoke.jl:44-52 that are implemented
LIMITERS = {
    "subcell": lambda: SubcellLimiter(bound=PositivityBound()),
    "zhangshu": lambda: ZhangShuLimiter(),
    "zhangshu_shock": lambda: ZhangShuLimiter(shockcapture=HennemannShockCapture()),
    "subcell_shock": lambda: SubcellLimiter(bound=PositivityBound(), shockcapture=HennemannShockCapture()),
    "minentropy": lambda: SubcellLimiter(bound=PositivityAndMinEntropyBound()),
    "relaxed_minentropy": lambda: SubcellLimiter(bound=PositivityAndRelaxedMinEntropyBound()),
    "cellentropy": lambda: SubcellLimiter(bound=PositivityAndCellEntropyBound()),
    "relaxed_cellentropy": lambda: SubcellLimiter(bound=PositivityAndRelaxedCellEntropyBound(beta=0.5)),
    # TVD density bounds (subcell.jl:87-141), not in the smoke test's list
    "tvd": lambda: SubcellLimiter(bound=TVDBound()),
    "tvd_minentropy": lambda: SubcellLimiter(bound=TVDAndMinEntropyBound()),
    "tvd_relaxed_minentropy": lambda: SubcellLimiter(bound=TVDAndRelaxedMinEntropyBound()),
    "tvd_cellentropy": lambda: SubcellLimiter(bound=TVDAndCellEntropyBound()),
    "tvd_relaxed_cellentropy": lambda: SubcellLimiter(bound=TVDAndRelaxedCellEntropyBound(beta=0.5)),
}

# dt0 = 0.9*dt_pos(U0)/CFL, recorded from the initial states (tests/golden/record_dt0.py)
RECORDED_DT0: Dict[str, float] = {
    "sod_1d": 0.0003169328455231937,
    "vortex_2d": 0.0009536460551185905,
    "riemann4_2d": 3.584113472465112e-05,
    "kelvin_helmholtz": 2.2148392215926948e-05,
}


def _param(N, K, xL, xR, T, CFL, dt0, Nd, zeta=0.1, eta=1.0, basis="lobatto", limiter="subcell", low_flux="nodal"):
    B, EPL = BASES[basis]
    return Param(N=N, K=K, xL=xL, xR=xR,
                 global_constants=GlobalConstant(POSTOL=1e-14, ZEROTOL=5e-16),
                 timestepping_param=TimesteppingParameter(T=T, CFL=CFL, dt0=dt0, t0=0.0),
                 limiting_param=LimitingParameter(zeta=zeta, eta=eta),
                 equation=CompressibleEulerIdealGas(Dim1() if Nd == 1 else Dim2(), 1.4),
                 rhs=ESLimitedLowOrderPos(low_order_surface_flux=(LaxFriedrichsOnProjectedVal() if low_flux == "projected"
                                                                  else LaxFriedrichsOnNodalVal())),
                 approximation_basis=B(),
                 entropyproj_limiter=EPL(),
                 rhs_limiter=LIMITERS[limiter]())


def _periodic_bc(param, md):
    md = make_periodic(md)
    return BCData(mapP=md.mapP), md


def _transmissive_bc(param, md):
    """Outflow (uP = own node, no LF dissipation) on every boundary face node
    (the mapO pattern of examples/2D/astro-jet.jl:53-55 and examples/1D/shu-osher.jl:23)."""
    Nfp = md.mapP.shape[0]
    fidx = np.flatnonzero(md.boundary.T.ravel()) + 1         # 1-based linear (Nfp, K) index
    return BCData(mapP=md.mapP, mapO=fidx.astype(np.int64))


# ------------------------------------------------------------------------------- Sod
def sod_ic(param, x):
    g = param.equation.gamma
    left = x < 0.5
    rho = np.where(left, 1.0, 0.125)
    u = np.zeros_like(x)
    p = np.where(left, 1.0, 0.1)
    return primitive_to_conservative(g, (rho, u, p))


def sod_1d(K=100, N=3, T=0.2, CFL=0.5, dt0=None, basis="lobatto", limiter="subcell", low_flux="nodal"):
    dt0 = RECORDED_DT0["sod_1d"] if dt0 is None else dt0
    return (_param(N, K, 0.0, 1.0, T, CFL, dt0, 1, basis=basis, limiter=limiter, low_flux=low_flux),
            sod_ic, _transmissive_bc)


# ------------------------------------------------------------------------------- Shu-Osher
def shu_osher_prim(x):
    """examples/1D/shu-osher.jl:8-14: the Mach 3 shock at x = -4 entering a density wave."""
    left = x < -4.0
    rho = np.where(left, 3.857143, 1 + 0.2 * np.sin(5 * x))
    u = np.where(left, 2.629369, 0.0)
    p = np.where(left, 10.3333, 1.0)
    return rho, u, p


def shu_osher_exact(equation, x, t):
    """The example's ``exact_sol`` (the initial state, as there: not an exact solution)."""
    return shu_osher_prim(x)


def shu_osher_ic(param, x):
    return primitive_to_conservative(param.equation.gamma, shu_osher_prim(x))


def _shu_osher_bc(param, md):
    """examples/1D/shu-osher.jl:16-31: inflow of the post-shock state at the left end
    (mapI = [1]), outflow at the right end (mapO = [2K])."""
    K = int(param.K)
    Ival = primitive_to_conservative(param.equation.gamma, (3.857143, 2.629369, 10.3333))
    return BCData(mapP=md.mapP, mapI=np.array([1], dtype=np.int64), mapO=np.array([2 * K], dtype=np.int64),
                  Ival=np.asarray(Ival, dtype=np.float64).reshape(1, 3))


def shu_osher(K=64, N=3, T=1.8, CFL=0.5, dt0=1e-4, basis="gauss", limiter="subcell"):
    """examples/1D/shu-osher.jl:43-55: [-5, 5], K = 64, N = 3, zeta 0.1, eta 1, Gauss +
    NodewiseScaledExtrapolation, SubcellLimiter(PositivityBound)."""
    return (_param(N, K, -5.0, 5.0, T, CFL, dt0, 1, zeta=0.1, eta=1.0, basis=basis, limiter=limiter),
            shu_osher_ic, _shu_osher_bc)


# ------------------------------------------------------------------------------- vortex
def vortex_exact_prim(x, y, t, gamma=1.4, beta=5.0, L=10.0, xc=0.0, yc=0.0):
    """Isentropic vortex advected with (1, 1) on a periodic box of side L (nearest image)."""
    dx = x - xc - t
    dy = y - yc - t
    dx = dx - L * np.round(dx / L)
    dy = dy - L * np.round(dy / L)
    r2 = dx * dx + dy * dy
    e = np.exp(1.0 - r2)
    u = 1.0 - beta * e * dy / (2.0 * math.pi)
    v = 1.0 + beta * e * dx / (2.0 * math.pi)
    rho = (1.0 - (gamma - 1.0) * beta * beta * np.exp(2.0 * (1.0 - r2)) / (16.0 * gamma * math.pi ** 2)) ** (1.0 / (gamma - 1.0))
    p = rho ** gamma
    return rho, u, v, p


def vortex_exact(equation, x, y, t):
    """``exact_sol(equation, x, y, t)`` form, as the reference's examples pass it to calculate_error."""
    return vortex_exact_prim(x, y, t, equation.gamma)


def vortex_ic(param, x, y):
    t0 = param.timestepping_param.t0
    return primitive_to_conservative(param.equation.gamma, vortex_exact_prim(x, y, t0, param.equation.gamma))


def vortex_2d(K=64, N=3, T=1.0, CFL=0.5, dt0=None, basis="lobatto", limiter="subcell"):
    dt0 = RECORDED_DT0["vortex_2d"] if dt0 is None else dt0
    return (_param(N, (K, K), (-5.0, -5.0), (5.0, 5.0), T, CFL, dt0, 2, basis=basis, limiter=limiter),
            vortex_ic, _periodic_bc)


# ------------------------------------------------------------------------------- 4-state Riemann
RIEMANN4_STATES = {  # (rho, u, v, p) by quadrant
    "NE": (1.5, 0.0, 0.0, 1.5),
    "NW": (0.5323, 1.206, 0.0, 0.3),
    "SW": (0.138, 1.206, 1.206, 0.029),
    "SE": (0.5323, 0.0, 1.206, 0.3),
}


def riemann4_ic(param, x, y):
    right = x > 0.5
    top = y > 0.5
    out = [np.zeros_like(x) for _ in range(4)]
    for key, mask in (("NE", right & top), ("NW", ~right & top), ("SW", ~right & ~top), ("SE", right & ~top)):
        for c in range(4):
            out[c] = np.where(mask, RIEMANN4_STATES[key][c], out[c])
    return primitive_to_conservative(param.equation.gamma, tuple(out))


def riemann4_2d(K=256, N=3, T=0.25, CFL=0.5, dt0=None, basis="lobatto", limiter="subcell", low_flux="nodal"):
    dt0 = RECORDED_DT0["riemann4_2d"] if dt0 is None else dt0
    return (_param(N, (K, K), (0.0, 0.0), (1.0, 1.0), T, CFL, dt0, 2, basis=basis, limiter=limiter,
                   low_flux=low_flux),
            riemann4_ic, _transmissive_bc)


# ------------------------------------------------------------------------------- Kelvin-Helmholtz
def kh_ic(param, x, y):
    B = np.tanh(15.0 * y + 7.5) - np.tanh(15.0 * y - 7.5)
    rho = 0.5 + 0.75 * B
    u = 0.5 * (B - 1.0)
    v = 0.1 * np.sin(2.0 * math.pi * x)
    p = np.ones_like(x)
    return primitive_to_conservative(param.equation.gamma, (rho, u, v, p))


def kelvin_helmholtz(K=512, N=4, T=10.0, CFL=0.5, dt0=None, basis="lobatto"):
    dt0 = RECORDED_DT0["kelvin_helmholtz"] if dt0 is None else dt0
    return _param(N, (K, K), (-1.0, -1.0), (1.0, 1.0), T, CFL, dt0, 2, basis=basis), kh_ic, _periodic_bc


def kelvin_helmholtz_gauss(K=80, N=3, T=10.0, CFL=0.5, dt0=1e-3, basis="gauss"):
    """The reference's example as written (examples/2D/kelvin-helmholtz.jl:43-55): 80x80, N=3,
    GaussCollocation + NodewiseScaledExtrapolation, zeta 0.1, eta 1.0, CFL 0.5, dt0 1e-3."""
    return _param(N, (K, K), (-1.0, -1.0), (1.0, 1.0), T, CFL, dt0, 2, basis=basis), kh_ic, _periodic_bc


# ------------------------------------------------------------------------------- astrophysical jet
def astro_jet_ic(param, x, y):
    g = param.equation.gamma
    return primitive_to_conservative(g, (np.full_like(x, 0.5), np.zeros_like(x), np.zeros_like(x),
                                         np.full_like(x, 0.4127)))


def _astro_jet_bc(param, md):
    """examples/2D/astro-jet.jl:20-59: periodic connectivity, then on the left boundary an inflow
    state (the jet rho 5, u 800 where |y| < 0.05, rest state elsewhere) and outflow on the right."""
    md = make_periodic(md)
    g = param.equation.gamma
    ZEROTOL = param.global_constants.ZEROTOL
    Kx, Ky = int(param.K[0]), int(param.K[1])
    N1D = param.N + 1
    Nfp = 4 * N1D
    K = Kx * Ky
    VX = -1.0 + 2.0 * np.arange(Kx + 1) / Kx
    VX[0], VX[-1] = -1.0, 1.0
    VX = (VX + 1.0) / 2.0 * (param.xR[0] - param.xL[0]) + param.xL[0]
    VY = -1.0 + 2.0 * np.arange(Ky + 1) / Ky
    VY[0], VY[-1] = -1.0, 1.0
    VY = (VY + 1.0) / 2.0 * (param.xR[1] - param.xL[1]) + param.xL[1]
    r1D = md_face_points(param)
    jet = primitive_to_conservative(g, (5.0, 800.0, 0.0, 0.4127))
    rest = primitive_to_conservative(g, (0.5, 0.0, 0.0, 0.4127))
    mapI, mapO, Ival = [], [], []
    for k in range(K):
        ex, ey = k % Kx, k // Kx
        for i in range(Nfp):
            face, m = divmod(i, N1D)
            rf, sf = [(-1.0, r1D[m]), (1.0, r1D[m]), (r1D[m], -1.0), (r1D[m], 1.0)][face]
            xf = 0.5 * (1.0 - rf) * VX[ex] + 0.5 * (1.0 + rf) * VX[ex + 1]
            yf = 0.5 * (1.0 - sf) * VY[ey] + 0.5 * (1.0 + sf) * VY[ey + 1]
            idx = i + Nfp * k + 1
            if abs(xf - param.xL[0]) < ZEROTOL:
                mapI.append(idx)
                Ival.append(jet if abs(yf) < 0.05 + ZEROTOL else rest)
            elif abs(xf - param.xR[0]) < ZEROTOL:
                mapO.append(idx)
    return BCData(mapP=md.mapP, mapI=np.array(mapI, dtype=np.int64), mapO=np.array(mapO, dtype=np.int64),
                  Ival=np.array(Ival, dtype=np.float64).reshape(-1, 4)), md


def md_face_points(param):
    """1D face quadrature points of the basis (LGL or Gauss nodes)."""
    from .operators import gauss_nodes_weights, lgl_nodes_weights
    if isinstance(param.approximation_basis, GaussCollocation):
        return gauss_nodes_weights(param.N)[0]
    return lgl_nodes_weights(param.N)[0]


def astro_jet(K=100, N=3, T=1e-3, CFL=0.5, dt0=1e-6, basis="gauss", limiter="relaxed_minentropy"):
    """The reference's astrophysical jet (examples/2D/astro-jet.jl:61-75): [-0.5,0.5]^2, 100x100,
    N=3, gamma 5/3, zeta 0.5, eta 0.1, Gauss + NodewiseScaledExtrapolation,
    SubcellLimiter(PositivityAndRelaxedMinEntropyBound), inflow left, outflow right."""
    param = _param(N, (K, K), (-0.5, -0.5), (0.5, 0.5), T, CFL, dt0, 2, zeta=0.5, eta=0.1, basis=basis,
                   limiter=limiter)
    param = replace(param, equation=CompressibleEulerIdealGas(Dim2(), 5.0 / 3.0))
    return param, astro_jet_ic, _astro_jet_bc


# ------------------------------------------------------------------------------- reference smoke
def smoke_exact_prim(x, y, t, gamma=1.4):
    """test/test_smoke.jl:6-20."""
    x0, y0, beta = 4.5, 5.0, 8.5
    r2 = (x - x0 - t) ** 2 + (y - y0) ** 2
    u = 1 - beta * np.exp(1 - r2) * (y - y0) / (2 * math.pi)
    v = beta * np.exp(1 - r2) * (x - x0 - t) / (2 * math.pi)
    rho = 1 - (1 / (8 * gamma * math.pi ** 2)) * (gamma - 1) / 2 * (beta * np.exp(1 - r2)) ** 2
    rho = rho ** (1 / (gamma - 1))
    p = rho ** gamma
    return rho, u, v, p


def smoke_exact(equation, x, y, t):
    """``exact_sol(eqn, x, y, t)`` of test/test_smoke.jl:6-20."""
    return smoke_exact_prim(x, y, t, equation.gamma)


def smoke_ic(param, x, y):
    return primitive_to_conservative(param.equation.gamma, smoke_exact_prim(x, y, param.timestepping_param.t0))


def smoke_vortex(N=3, K=5, T=2e-2, CFL=1.0, dt0=1e-2, basis="lobatto", limiter="subcell", low_flux="nodal"):
    """test/test_smoke.jl:54-66,117-129; ``limiter`` picks one of its limiter set-ups."""
    return (_param(N, (K, K), (0.0, 0.0), (10.0, 10.0), T, CFL, dt0, 2, zeta=0.1, eta=0.5, basis=basis,
                   limiter=limiter, low_flux=low_flux),
            smoke_ic, _periodic_bc)


# ------------------------------------------------------------------------------- batched workload
MASK64 = (1 << 64) - 1


class SplitMix64:
    """splitmix64 (Steele, Lea & Flood 2014): the seeded stream the batched workload is drawn from."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def integer(self, lo: int, hi: int) -> int:
        """uniform in {lo, ..., hi}"""
        return lo + self.next() % (hi - lo + 1)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next() >> 11) * 2.0 ** -53)


def periodic_wave_ic(param, x, y, kx=1, ky=1, u0=0.0, v0=0.0, dip=None, sigma=0.05):
    """rho = 1 + 0.5 sin(2 pi kx x) sin(2 pi ky y), velocity (u0, v0), p = 1; with ``dip = (x0, y0)``
    the density is blended towards 1e-6 with the Gaussian weight g = exp(-|x - x0|^2 / (2 sigma^2))
    (nearest periodic image): rho - (rho - 1e-6) g, which is 1e-6 at x0 and positive everywhere
    (subtracting the fixed amount (rho(x0) - 1e-6) g instead leaves negative densities where
    the background falls faster than the Gaussian)."""
    g = param.equation.gamma
    rho = 1.0 + 0.5 * np.sin(2 * math.pi * kx * x) * np.sin(2 * math.pi * ky * y)
    if dip is not None:
        x0, y0 = dip
        dx = x - x0 - np.round(x - x0)
        dy = y - y0 - np.round(y - y0)
        w = np.exp(-(dx * dx + dy * dy) / (2 * sigma * sigma))
        rho = rho - (rho - 1e-6) * w
    return primitive_to_conservative(g, (rho, np.full_like(x, u0), np.full_like(x, v0), np.ones_like(x)))


def periodic_wave(Kx=32, Ky=32, N=3, kx=1, ky=1, u0=0.0, v0=0.0, dip=None, T=1.0, CFL=0.5, dt0=1e-4,
                  basis="lobatto"):
    param = _param(N, (Kx, Ky), (0.0, 0.0), (1.0, 1.0), T, CFL, dt0, 2, basis=basis)
    return param, (lambda prm, x, y: periodic_wave_ic(prm, x, y, kx, ky, u0, v0, dip)), _periodic_bc


def batched_draws(count: int = 48, seed: int = 0):
    """The instance parameters of the batched workload, drawn in this order per instance:
    Kx, Ky in {32..160}; N in {2, 3, 4}; kx, ky in {1..4}; u0, v0 in [-1, 1]; a dip with
    probability 0.1 at (x0, y0) uniform in [0, 1)^2."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        d = dict(Kx=rng.integer(32, 160), Ky=rng.integer(32, 160), N=rng.integer(2, 4),
                 kx=rng.integer(1, 4), ky=rng.integer(1, 4), u0=rng.uniform(-1.0, 1.0), v0=rng.uniform(-1.0, 1.0))
        if rng.uniform() < 0.1:
            d["dip"] = (rng.uniform(), rng.uniform())
        out.append(d)
    return out


def batched_instances(count: int = 48, seed: int = 0, T: float = 1.0, CFL: float = 0.5, scale: float = 1.0):
    """``(solver, state, state_param)`` of every instance of the batched workload, each with
    dt0 = 0.9 * dt_pos(U0) / CFL from the host CFL estimate.  ``scale`` shrinks the element
    counts (for quick tests; 1.0 is the workload)."""
    from .dg import flat_problem
    from .timestep import cfl_dt
    out = []
    for d in batched_draws(count, seed):
        kw = dict(d)
        kw["Kx"] = max(2, int(round(kw["Kx"] * scale)))
        kw["Ky"] = max(2, int(round(kw["Ky"] * scale)))
        param, ic, bc = periodic_wave(T=T, CFL=CFL, **kw)
        solver, state, sp = initialize_DG(param, ic, bc)
        dt = cfl_dt(flat_problem(solver, sp), state.preallocation.Uq)
        param = replace(param, timestepping_param=replace(param.timestepping_param, dt0=0.9 * dt / CFL))
        solver = replace(solver, param=param)
        out.append((solver, state, sp))
    return out


CONFIGS: Dict[str, Callable] = {
    "sod_1d": sod_1d,
    "vortex_2d": vortex_2d,
    "riemann4_2d": riemann4_2d,
    "kelvin_helmholtz": kelvin_helmholtz,
    "kelvin_helmholtz_gauss": kelvin_helmholtz_gauss,
    "astro_jet": astro_jet,
    "smoke_vortex": smoke_vortex,
    "shu_osher": shu_osher,
    "periodic_wave": periodic_wave,
}


def build(name: str, **kw):
    """Return ``(solver, state, state_param)`` for a named problem."""
    param, ic, bc = CONFIGS[name](**kw)
    return initialize_DG(param, ic, bc)
