"""Option and parameter types of the solver, mirroring the reference's dispatch types.

The reference selects methods by the type parameters of ``Param``
(src/common/types/Solver.jl:151-170).  Here each option is a small frozen
dataclass with the same name; ``Param`` carries them, and the native library
receives them as integer enums (``include/p2de_hip.h``).  Combinations the
MI355X path does not implement are rejected with ``UnsupportedOption`` before
anything reaches the device.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Tuple, Union


class UnsupportedOption(ValueError):
    """Raised for an option combination the HIP path does not implement."""


# --- dimensions and equations (Solver.jl:97-121) ---------------------------------
@dataclass(frozen=True)
class Dim1:
    pass


@dataclass(frozen=True)
class Dim2:
    pass


@dataclass(frozen=True)
class CompressibleEulerIdealGas:
    """``CompressibleEulerIdealGas{DIM}(gamma)`` (Solver.jl:110-118)."""
    dim: Union[Dim1, Dim2]
    gamma: float = 1.4

    @property
    def Nd(self) -> int:
        return 1 if isinstance(self.dim, Dim1) else 2

    @property
    def Nc(self) -> int:  # num_components (traits.jl:408-414)
        return self.Nd + 2


# --- fluxes (Solver.jl:15-36) ----------------------------------------------------
@dataclass(frozen=True)
class ChandrashekarFlux:
    pass


@dataclass(frozen=True)
class CentralFlux:
    pass


@dataclass(frozen=True)
class LaxFriedrichsOnNodalVal:
    pass


@dataclass(frozen=True)
class LaxFriedrichsOnProjectedVal:
    pass


@dataclass(frozen=True)
class ChandrashekarOnProjectedVal:
    pass


@dataclass(frozen=True)
class LimitedDG:
    """``LimitedDG`` / ``ESLimitedLowOrderPos`` (Solver.jl:9-13,26,31-33)."""
    low_order_surface_flux: object = field(default_factory=LaxFriedrichsOnNodalVal)
    high_order_surface_flux: object = field(default_factory=LaxFriedrichsOnProjectedVal)
    high_order_volume_flux: object = field(default_factory=ChandrashekarFlux)


def ESLimitedLowOrderPos(low_order_surface_flux=None, high_order_surface_flux=None) -> LimitedDG:
    """Constructor alias of Solver.jl:31-33."""
    return LimitedDG(low_order_surface_flux or LaxFriedrichsOnNodalVal(),
                     high_order_surface_flux or LaxFriedrichsOnProjectedVal(),
                     ChandrashekarFlux())


# --- limiters (Solver.jl:38-87) --------------------------------------------------
@dataclass(frozen=True)
class NoEntropyProjectionLimiter:
    pass


@dataclass(frozen=True)
class NodewiseScaledExtrapolation:
    pass


@dataclass(frozen=True)
class PositivityBound:
    pass


@dataclass(frozen=True)
class PositivityAndMinEntropyBound:
    pass


@dataclass(frozen=True)
class PositivityAndRelaxedMinEntropyBound:
    pass


@dataclass(frozen=True)
class PositivityAndCellEntropyBound:
    pass


@dataclass(frozen=True)
class PositivityAndRelaxedCellEntropyBound:
    beta: float = 0.5


@dataclass(frozen=True)
class TVDBound:
    pass


@dataclass(frozen=True)
class TVDAndMinEntropyBound:
    pass


@dataclass(frozen=True)
class TVDAndRelaxedMinEntropyBound:
    pass


@dataclass(frozen=True)
class TVDAndCellEntropyBound:
    pass


@dataclass(frozen=True)
class TVDAndRelaxedCellEntropyBound:
    beta: float = 0.5


@dataclass(frozen=True)
class NoShockCapture:
    pass


@dataclass(frozen=True)
class HennemannShockCapture:
    a: float = 0.5
    c: float = 1.8


@dataclass(frozen=True)
class SubcellLimiter:
    bound: object = field(default_factory=PositivityBound)
    shockcapture: object = field(default_factory=NoShockCapture)


@dataclass(frozen=True)
class ZhangShuLimiter:
    shockcapture: object = field(default_factory=NoShockCapture)


# --- bases (Solver.jl:89-95) -----------------------------------------------------
@dataclass(frozen=True)
class LobattoCollocation:
    pass


@dataclass(frozen=True)
class GaussCollocation:
    pass


# --- scalar parameter groups (Solver.jl:123-149) ---------------------------------
@dataclass(frozen=True)
class GlobalConstant:
    POSTOL: float = 1e-14
    ZEROTOL: float = 5e-16


@dataclass(frozen=True)
class TimesteppingParameter:
    T: float
    CFL: float
    dt0: float
    t0: float = 0.0


@dataclass(frozen=True)
class LimitingParameter:
    zeta: float = 0.1
    eta: float = 1.0


@dataclass(frozen=True)
class PostprocessingParameter:
    output_interval: int = 10**9


@dataclass(frozen=True)
class Param:
    """``Param`` (Solver.jl:151-170).  ``K`` is an int in 1D and ``(Kx, Ky)`` in 2D;
    ``xL``/``xR`` are floats in 1D and pairs in 2D, as in the reference."""
    N: int
    K: Union[int, Tuple[int, int]]
    xL: Union[float, Tuple[float, float]]
    xR: Union[float, Tuple[float, float]]
    global_constants: GlobalConstant
    timestepping_param: TimesteppingParameter
    limiting_param: LimitingParameter
    equation: CompressibleEulerIdealGas
    rhs: LimitedDG = field(default_factory=LimitedDG)
    approximation_basis: object = field(default_factory=LobattoCollocation)
    entropyproj_limiter: object = field(default_factory=NoEntropyProjectionLimiter)
    rhs_limiter: object = field(default_factory=SubcellLimiter)
    postprocessing_param: PostprocessingParameter = field(default_factory=PostprocessingParameter)

    @property
    def num_elements(self) -> int:  # traits.jl:424-440
        if self.equation.Nd == 1:
            return int(self.K)
        return int(self.K[0]) * int(self.K[1])


# SubcellLimiter bounds in the order of their C-ABI codes (P2DE_*_BOUND): positivity or TVD density
# bounds, each with no entropy bound, the min-entropy bound, its relaxed form, the cell-entropy
# bound or its relaxed form
BOUNDS = (PositivityBound, PositivityAndMinEntropyBound, PositivityAndRelaxedMinEntropyBound,
          PositivityAndCellEntropyBound, PositivityAndRelaxedCellEntropyBound,
          TVDBound, TVDAndMinEntropyBound, TVDAndRelaxedMinEntropyBound,
          TVDAndCellEntropyBound, TVDAndRelaxedCellEntropyBound)
CELL_ENTROPY_BOUNDS = (PositivityAndCellEntropyBound, PositivityAndRelaxedCellEntropyBound,
                       TVDAndCellEntropyBound, TVDAndRelaxedCellEntropyBound)


def limiter_codes(lim) -> tuple:
    """(rhs_limiter, bound, shockcapture, a, c, cell_beta) as the C-ABI enums take them."""
    rl = 1 if isinstance(lim, ZhangShuLimiter) else 0
    bd = 0
    beta = 0.5
    if isinstance(lim, SubcellLimiter):
        bd = BOUNDS.index(type(lim.bound)) if type(lim.bound) in BOUNDS else -1
        if isinstance(lim.bound, (PositivityAndRelaxedCellEntropyBound, TVDAndRelaxedCellEntropyBound)):
            beta = float(lim.bound.beta)
    sc = lim.shockcapture
    if isinstance(sc, HennemannShockCapture):
        return rl, bd, 1, float(sc.a), float(sc.c), beta
    return rl, bd, 0, 0.5, 1.8, beta


def check_supported(param: Param) -> None:
    """Reject option combinations outside the implemented hot path.

    Implemented: ``ESLimitedLowOrderPos(LaxFriedrichsOnNodalVal, LaxFriedrichsOnProjectedVal)``
    + ``SubcellLimiter(PositivityBound, NoShockCapture)`` for 1D/2D Euler with
    ``LobattoCollocation`` + ``NoEntropyProjectionLimiter`` (the combination
    test/test_smoke.jl:107-108 exercises) or ``GaussCollocation`` with
    ``NodewiseScaledExtrapolation`` (examples/2D/kelvin-helmholtz.jl:50-55) or
    ``NoEntropyProjectionLimiter``; the limiter may also be ``ZhangShuLimiter``, the bound
    ``PositivityAndMinEntropyBound`` or ``PositivityAndRelaxedMinEntropyBound``, and either
    limiter may use ``HennemannShockCapture`` (the first five limiter set-ups of
    test/test_smoke.jl:44-52).
    """
    rhs = param.rhs
    if not isinstance(rhs, LimitedDG):
        raise UnsupportedOption(f"rhs {rhs!r}: only LimitedDG (ESLimitedLowOrderPos) is implemented")
    if not isinstance(rhs.high_order_volume_flux, ChandrashekarFlux):
        raise UnsupportedOption("high_order_volume_flux must be ChandrashekarFlux")
    if not isinstance(rhs.low_order_surface_flux, (LaxFriedrichsOnNodalVal, LaxFriedrichsOnProjectedVal)):
        raise UnsupportedOption("low_order_surface_flux: LaxFriedrichsOnNodalVal or LaxFriedrichsOnProjectedVal")
    if not isinstance(rhs.high_order_surface_flux, LaxFriedrichsOnProjectedVal):
        raise UnsupportedOption("high_order_surface_flux must be LaxFriedrichsOnProjectedVal")
    basis = param.approximation_basis
    if not isinstance(basis, (LobattoCollocation, GaussCollocation)):
        raise UnsupportedOption(f"approximation_basis {basis!r}: LobattoCollocation or GaussCollocation")
    epl = param.entropyproj_limiter
    if isinstance(basis, LobattoCollocation) and not isinstance(epl, NoEntropyProjectionLimiter):
        # the Lobatto method of compute_entropyproj_limiting_param! (filter.jl:1-3) does not run
        raise UnsupportedOption("LobattoCollocation takes NoEntropyProjectionLimiter")
    if not isinstance(epl, (NoEntropyProjectionLimiter, NodewiseScaledExtrapolation)):
        raise UnsupportedOption(f"entropyproj_limiter {epl!r}: NoEntropyProjectionLimiter or "
                                "NodewiseScaledExtrapolation")
    lim = param.rhs_limiter
    if not isinstance(lim, (SubcellLimiter, ZhangShuLimiter)):
        raise UnsupportedOption(f"rhs_limiter {lim!r}: SubcellLimiter or ZhangShuLimiter")
    if isinstance(lim, SubcellLimiter) and not isinstance(lim.bound, BOUNDS):
        raise UnsupportedOption(f"SubcellLimiter bound {lim.bound!r}: " + ", ".join(b.__name__ for b in BOUNDS))
    if isinstance(lim, SubcellLimiter) and isinstance(lim.bound, CELL_ENTROPY_BOUNDS):
        # 1D: the reference sizes the bound's scratch (Nq - N1D) x K = 0 x K and indexes past it
        # (State.jl:182-183, subcell.jl:503-505); Gauss: its interface step (subcell.jl:745-789)
        # reads neighbour parameters the neighbour's own step may already have changed
        if param.equation.Nd != 2 or not isinstance(basis, LobattoCollocation):
            raise UnsupportedOption("the cell-entropy bounds are implemented for 2D LobattoCollocation")
    if not isinstance(lim.shockcapture, (NoShockCapture, HennemannShockCapture)):
        raise UnsupportedOption(f"shockcapture {lim.shockcapture!r}: NoShockCapture or HennemannShockCapture")
    if not 1 <= param.N <= 4:
        raise UnsupportedOption(f"polynomial degree N={param.N}: 1..4 are implemented")
