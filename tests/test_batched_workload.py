"""The batched workload's generator: splitmix64 against its published outputs, the draws'
ranges, and the initial states (positive everywhere, near-vacuum dips reaching 1e-6)."""
import math

import numpy as np

from conftest import PKG


def test_splitmix64_known_outputs():
    # first outputs of splitmix64 seeded with 0 (Steele, Lea & Flood 2014; the reference
    # implementation by Vigna, xoshiro.di.unimi.it/splitmix64.c)
    r = PKG.configs.SplitMix64(0)
    assert [r.next() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    r = PKG.configs.SplitMix64(0)
    u = [r.uniform() for _ in range(1000)]
    assert min(u) >= 0.0 and max(u) < 1.0 and abs(np.mean(u) - 0.5) < 0.05


def test_draws_ranges_and_determinism():
    d = PKG.configs.batched_draws()
    assert d == PKG.configs.batched_draws() and len(d) == 48
    for x in d:
        assert 32 <= x["Kx"] <= 160 and 32 <= x["Ky"] <= 160 and x["N"] in (2, 3, 4)
        assert 1 <= x["kx"] <= 4 and 1 <= x["ky"] <= 4 and -1 <= x["u0"] <= 1 and -1 <= x["v0"] <= 1
    ndip = sum("dip" in x for x in d)
    assert 1 <= ndip <= 12          # probability 0.1 per instance
    dofs = sum(x["Kx"] * x["Ky"] * (x["N"] + 1) ** 2 for x in d)
    assert 6e6 < dofs < 9e6         # about 7.4M expected


def test_dip_initial_condition_positive_and_near_vacuum():
    param, ic, bc = PKG.configs.periodic_wave(Kx=64, Ky=64, N=3, kx=4, ky=3, u0=0.2, v0=-0.3, dip=(0.37, 0.61))
    solver, state, sp = PKG.initialize_DG(param, ic, bc)
    U = state.preallocation.Uq
    rho = U[..., 0]
    p = 0.4 * (U[..., 3] - 0.5 * (U[..., 1] ** 2 + U[..., 2] ** 2) / rho)
    assert rho.min() > 0 and np.allclose(p, 1.0)
    # the density at the dip centre itself is 1e-6 (evaluate the initial condition there)
    x = np.array([[0.37]])
    y = np.array([[0.61]])
    c = ic(param, x, y)
    assert math.isclose(c[0, 0, 0], 1e-6, rel_tol=1e-9)
    # the background far from the dip is untouched
    x = np.array([[0.87]])
    y = np.array([[0.11]])
    c = ic(param, x, y)
    assert math.isclose(c[0, 0, 0], 1.0 + 0.5 * math.sin(2 * math.pi * 4 * 0.87) * math.sin(2 * math.pi * 3 * 0.11), rel_tol=1e-12)


def test_astro_jet_boundary_data():
    """examples/2D/astro-jet.jl:20-59: every left face node is inflow (the jet state where
    |y| < 0.05, the rest state elsewhere), every right face node outflow, the rest periodic."""
    s, st, sp = PKG.configs.build("astro_jet", K=20)
    bc = sp.bcdata
    N1D = s.param.N + 1
    assert len(bc.mapI) == len(bc.mapO) == 20 * N1D
    Ival = np.asarray(bc.Ival)
    jet = Ival[:, 0] == 5.0
    assert jet.sum() >= 1 and np.all(Ival[~jet, 0] == 0.5)
    assert np.allclose(Ival[jet, 1] / Ival[jet, 0], 800.0)
