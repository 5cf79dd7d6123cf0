"""The host CFL time step (p2de_jl_amd.timestep.cfl_dt), used to set dt0, against the
oracle's stage-1 dt (calculate_lambda_and_low_order_CFL!, low_order_graph_viscosity.jl:222-285)
with the CFL*dt0 and T caps lifted: equal to the last bit on 1D and 2D problems and perturbed
states (the numpy sums run in the same order for these operators)."""
import numpy as np
import pytest

from conftest import PKG, perturbed_state, problem
from oracle.oracle import Oracle


@pytest.mark.parametrize("name,kw", [("vortex_2d", dict(K=8, dt0=1e30)), ("vortex_2d", dict(K=5, N=4, dt0=1e30)),
                                     ("riemann4_2d", dict(K=12, dt0=1e30)), ("sod_1d", dict(dt0=1e30)),
                                     ("kelvin_helmholtz", dict(K=6, dt0=1e30))])
def test_cfl_dt_matches_oracle(name, kw):
    solver, state, sp, fp = problem(name, **kw)
    fp.T = 1e30
    orc = Oracle(fp)
    for seed, U in enumerate([state.preallocation.Uq, perturbed_state(state.preallocation.Uq, seed=5)]):
        a = PKG.timestep.cfl_dt(fp, U)
        b = orc.rhs(U, 0.0, 1e30, 1)
        assert np.isfinite(a) and a > 0
        assert abs(a - b) <= 1e-14 * b, (name, seed, a, b)
