"""Pins of the oracle's pointwise physics (CPU).

Each check is independent of the oracle's own formulas: mpmath values from textbook
definitions (tests/golden/pointwise.json, made by tests/golden/make_pointwise.py) and
mathematical identities a dropped term or a wrong sign would break:
  * v(u) equals the gradient of -rho*s (numerical differentiation at 50 digits);
  * u(v(u)) = u;
  * fS is symmetric, consistent (fS(u,u) = f(u)) and satisfies Tadmor's condition
    (v_L - v_R) . fS = psi_L - psi_R with psi = (gamma-1) rho u;
  * the log mean matches (b-a)/(ln b - ln a) away from the series branch and, in the branch,
    deviates by the amount the reference's coefficients imply (-0.2 instead of -1/12);
  * the limiter's quadratic root is the largest admissible step found by brute force.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

GOLD = json.loads((Path(__file__).parent / "golden" / "pointwise.json").read_text())
G = 1.4


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("case", range(len(GOLD["logmean"])))
def test_logmean_against_mpmath(case):
    g = GOLD["logmean"][case]
    a, b, lm = g["a"], g["b"], g["lm"]
    got = O.pt_logmean(a, b)
    f = (b - a) / (0.5 * (a + b))
    if abs(f) >= 1e-4:
        # the log difference cancels: error ~ ulp(log a) / |log b - log a|
        tol = 4e-16 * max(1.0, abs(math.log(a))) / abs(math.log(b) - math.log(a)) + 4e-16
        assert abs(got - lm) <= tol * abs(lm)
    else:
        # series branch aavg (1 - 0.2 f^2 + ...) vs the true 1 - f^2/12: deviation (0.2 - 1/12) f^2
        dev = (got - lm) / lm
        assert abs(dev + (0.2 - 1.0 / 12.0) * f * f) <= 1e-3 * (0.2 - 1 / 12) * f * f + 2e-16


def test_logmean_branch_switch_is_the_reference_series():
    # just inside the switch the reference series is 1.17e-9 below the log mean
    a = 1.0
    b = a * (1 + 0.99e-4)
    f = (b - a) / (0.5 * (a + b))
    got = O.pt_logmean(a, b)
    true = (b - a) / (math.log(b) - math.log(a))
    assert 1.0e-9 < (true - got) / true < 1.3e-9
    assert abs(f) < 1e-4


@pytest.mark.parametrize("case", range(len(GOLD["entropy_vars"])))
def test_entropy_variables_are_the_entropy_gradient(case):
    g = GOLD["entropy_vars"][case]
    V = O.pt_v_of_u(g["Nd"], G, g["U"])
    assert rel(V, g["V"]) <= 5e-13
    U = O.pt_u_of_v(g["Nd"], G, V)
    assert rel(U, g["U"]) <= 5e-13


@pytest.mark.parametrize("case", range(len(GOLD["euler_flux"])))
def test_euler_flux_textbook(case):
    g = GOLD["euler_flux"][case]
    F = O.pt_flux(g["Nd"], G, g["U"])
    assert rel(F, g["F"]) <= 1e-14


@pytest.mark.parametrize("case", range(len(GOLD["chandrashekar"])))
def test_chandrashekar_flux_published_form(case):
    g = GOLD["chandrashekar"][case]
    F = O.pt_fS(g["Nd"], G, g["UL"], g["UR"])
    assert rel(F, g["F"]) <= 1e-12


def _rand_state(rng, Nd, lo=1e-2, hi=10.0):
    rho = math.exp(rng.uniform(math.log(lo), math.log(hi)))
    p = math.exp(rng.uniform(math.log(lo), math.log(hi)))
    vel = rng.uniform(-2, 2, size=Nd)
    E = p / (G - 1) + 0.5 * rho * float(np.sum(vel * vel))
    return np.array([rho, *(rho * vel), E])


@pytest.mark.parametrize("Nd", [1, 2])
def test_fS_symmetry_consistency_tadmor(Nd):
    rng = np.random.default_rng(7)
    for _ in range(50):
        UL, UR = _rand_state(rng, Nd), _rand_state(rng, Nd)
        F = O.pt_fS(Nd, G, UL, UR)
        assert np.array_equal(F, O.pt_fS(Nd, G, UR, UL))        # symmetric, bit for bit
        assert rel(O.pt_fS(Nd, G, UL, UL), O.pt_flux(Nd, G, UL)) <= 1e-14
        vL, vR = O.pt_v_of_u(Nd, G, UL), O.pt_v_of_u(Nd, G, UR)
        for d in range(Nd):
            psiL, psiR = (G - 1) * UL[1 + d], (G - 1) * UR[1 + d]
            lhs = float(np.dot(vL - vR, F[d]))
            scale = float(np.sum(np.abs(vL - vR) * np.abs(F[d]))) + abs(psiL) + abs(psiR)
            assert abs(lhs - (psiL - psiR)) <= 1e-12 * scale


def test_wavespeed_is_normal_speed_plus_sound_speed_with_rotated_pressure():
    rng = np.random.default_rng(3)
    for _ in range(40):
        U = _rand_state(rng, 2)
        th = rng.uniform(0, 2 * math.pi)
        n = np.array([math.cos(th), math.sin(th)])
        rho = U[0]
        u = U[1:3] / rho
        p = (G - 1) * (U[3] - 0.5 * rho * float(u @ u))
        un = float(u @ n)
        ut2 = float(u @ u) - un * un
        # the reference rotates to a 1D state, so the tangential kinetic energy counts as
        # internal energy in the sound speed (compressible_Navier_Stokes.jl:58-62)
        c_rot = math.sqrt(G * (p + (G - 1) * 0.5 * rho * ut2) / rho)
        assert abs(O.pt_wavespeed(2, G, U, n) - (abs(un) + c_rot)) <= 1e-13 * (abs(un) + c_rot)
        # axis-aligned normal with no tangential velocity: the textbook |u.n| + c
        U2 = U.copy()
        U2[2] = 0.0
        U2[3] = p / (G - 1) + 0.5 * rho * u[0] ** 2
        c = math.sqrt(G * p / rho)
        assert abs(O.pt_wavespeed(2, G, U2, np.array([1.0, 0.0])) - (abs(u[0]) + c)) <= 1e-13 * (abs(u[0]) + c)


def _rhoe(U):
    return U[-1] - 0.5 * float(np.sum(U[1:-1] ** 2)) / U[0]


@pytest.mark.parametrize("Nd", [1, 2])
def test_limiting_param_matches_brute_force(Nd):
    """Intended bounds (Urho = Inf): l is the largest step in [0,1] keeping
    rho >= zeta rho_L and rhoe >= zeta rhoe_L, found by a fine scan (the quadratic and
    linear constraints are monotone-crossing on [0, l*])."""
    rng = np.random.default_rng(11)
    zeta = 0.1
    checked = 0
    for _ in range(400):
        uL = _rand_state(rng, Nd)
        P = rng.normal(size=Nd + 2) * uL * rng.uniform(0.1, 3.0)
        l = O.pt_limiting_param(Nd, zeta, 5e-16, 1, uL, P)
        assert 0.0 <= l <= 1.0
        ls = np.linspace(0, 1, 20001)
        W = uL[None, :] + ls[:, None] * P[None, :]
        rhoe_W = W[:, -1] - 0.5 * np.sum(W[:, 1:-1] ** 2, axis=1) / W[:, 0]
        ok = (W[:, 0] >= zeta * uL[0] - 1e-12) & (rhoe_W >= zeta * _rhoe(uL) - 1e-12 * abs(_rhoe(uL)))
        first_bad = int(np.argmax(~ok)) if not ok.all() else len(ls)
        l_brute = ls[first_bad - 1] if first_bad < len(ls) else 1.0
        assert abs(l - l_brute) <= 1e-4 + 1e-9
        checked += 1
    assert checked == 400


def test_limiting_param_reference_bounds_quirk():
    """Reference bounds: limiter_utils.jl:11 reads Urho from the tuple's third slot
    (Lphi = 0), so any step with P_rho >= +0 and rho_L + P_rho > 0 gets l = 0."""
    uL = np.array([1.0, 0.1, 0.0, 2.5])
    P = np.array([1e-3, 0.0, 0.0, 0.0])
    assert O.pt_limiting_param(2, 0.1, 5e-16, 0, uL, P) == 0.0
    assert O.pt_limiting_param(2, 0.1, 5e-16, 1, uL, P) == 1.0
    P = np.array([-1e-3, 0.0, 0.0, 0.0])      # negative density increment: not limited by Urho
    assert O.pt_limiting_param(2, 0.1, 5e-16, 0, uL, P) == 1.0
