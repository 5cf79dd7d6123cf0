"""Generate tests/golden/pointwise.json: high-precision (mpmath, 50 digits) values that
pin the oracle's pointwise physics independently of its own formulas.

* LGL nodes/weights: roots of (1 - x^2) P_N'(x), weights 2/(N(N+1) P_N(x)^2).
* log mean: (b - a) / (ln b - ln a).
* entropy variables: the gradient of U(u) = -rho * s(u), s = ln(p / rho^gamma), by
  numerical differentiation at 50 digits (not the closed form the solver uses).
* Euler flux: the textbook physical flux.
* Chandrashekar EC flux in the form published by Chandrashekar (2013),
  F_E = F_rho / (2 (gamma-1) beta_ln) - 1/2 avg(|u|^2) F_rho + avg(u) . F_m,
  with the true log means (states chosen with |da/a| > 1e-3, outside the series branch).

Run from the repository root:  python tests/golden/make_pointwise.py
"""
import json
from pathlib import Path

import mpmath as mp
import numpy as np

mp.mp.dps = 50
OUT = Path(__file__).resolve().parent / "pointwise.json"
G = mp.mpf("1.4")


def lgl(N):
    x = mp.mpf
    P = lambda n, t: mp.legendre(n, t)
    dP = lambda t: mp.diff(lambda s: P(N, s), t)
    # interior roots of P_N' from Chebyshev-Gauss-Lobatto guesses
    nodes = [x(-1)]
    for i in range(1, N):
        guess = -mp.cos(mp.pi * i / N)
        nodes.append(mp.findroot(dP, guess))
    nodes.append(x(1))
    w = [2 / (N * (N + 1) * P(N, t) ** 2) for t in nodes]
    return [float(t) for t in nodes], [float(t) for t in w]


def pressure(U):
    if len(U) == 3:
        rho, m, E = U
        return (G - 1) * (E - m * m / (2 * rho))
    rho, m1, m2, E = U
    return (G - 1) * (E - (m1 * m1 + m2 * m2) / (2 * rho))


def entropy_U(U):
    rho = U[0]
    s = mp.log(pressure(U) / rho ** G)
    return -rho * s


def entropy_vars(U):
    U = [mp.mpf(u) for u in U]
    out = []
    for c in range(len(U)):
        def f(t, c=c):
            V = list(U)
            V[c] = t
            return entropy_U(V)
        out.append(mp.diff(f, U[c]))
    return out


def euler_flux(U):
    p = pressure(U)
    if len(U) == 3:
        rho, m, E = U
        u = m / rho
        return [[m, m * u + p, u * (E + p)]]
    rho, m1, m2, E = U
    u, v = m1 / rho, m2 / rho
    return [[m1, m1 * u + p, m1 * v, u * (E + p)], [m2, m2 * u, m2 * v + p, v * (E + p)]]


def logmean(a, b):
    a, b = mp.mpf(a), mp.mpf(b)
    return (b - a) / (mp.log(b) - mp.log(a))


def chandrashekar(UL, UR):
    UL = [mp.mpf(u) for u in UL]
    UR = [mp.mpf(u) for u in UR]
    Nd = len(UL) - 2
    betaL = UL[0] / (2 * pressure(UL))
    betaR = UR[0] / (2 * pressure(UR))
    rho_ln = logmean(UL[0], UR[0])
    beta_ln = logmean(betaL, betaR)
    velL = [UL[1 + d] / UL[0] for d in range(Nd)]
    velR = [UR[1 + d] / UR[0] for d in range(Nd)]
    avg = [(velL[d] + velR[d]) / 2 for d in range(Nd)]
    p_hat = (UL[0] + UR[0]) / 2 / (betaL + betaR)
    avg_sq = sum((velL[d] ** 2 + velR[d] ** 2) / 2 for d in range(Nd))
    F = []
    for d in range(Nd):
        Frho = rho_ln * avg[d]
        Fm = [Frho * avg[e] + (p_hat if e == d else 0) for e in range(Nd)]
        FE = Frho / (2 * (G - 1) * beta_ln) - avg_sq / 2 * Frho + sum(avg[e] * Fm[e] for e in range(Nd))
        F.append([Frho] + Fm + [FE])
    return F


def prim2cons(rho, vel, p):
    rho, p = mp.mpf(rho), mp.mpf(p)
    vel = [mp.mpf(v) for v in vel]
    E = p / (G - 1) + rho * sum(v * v for v in vel) / 2
    return [rho] + [rho * v for v in vel] + [E]


def f2(x):
    if isinstance(x, list):
        return [f2(y) for y in x]
    return float(x)


def main():
    rng = np.random.default_rng(20261202)
    data = {"lgl": {}, "logmean": [], "entropy_vars": [], "euler_flux": [], "chandrashekar": []}
    for N in range(1, 5):
        x, w = lgl(N)
        data["lgl"][str(N)] = {"x": x, "w": w}
    # log mean over magnitudes, including pairs inside the |f| < 1e-4 series branch
    for a, rel in [(1.0, 0.5), (0.2, 3.0), (1e-6, 0.3), (7.0, 1e-2), (3.0, 1e-3), (1.0, 2e-4),
                   (2.5, 9.9e-5), (0.8, 5e-5), (1.3, 1e-6), (4.0, 1e-9)]:
        b = a * (1 + rel)
        data["logmean"].append({"a": a, "b": b, "lm": float(logmean(a, b))})
    for Nd in (1, 2):
        for trial in range(6):
            rho = float(np.exp(rng.uniform(np.log(1e-3), np.log(10.0))))
            p = float(np.exp(rng.uniform(np.log(1e-3), np.log(10.0))))
            vel = [float(v) for v in rng.uniform(-2, 2, size=Nd)]
            U = prim2cons(rho, vel, p)
            Uf = [float(u) for u in U]
            U = [mp.mpf(u) for u in Uf]   # the exact double inputs
            data["entropy_vars"].append({"Nd": Nd, "U": Uf, "V": f2(entropy_vars(U))})
            data["euler_flux"].append({"Nd": Nd, "U": Uf, "F": f2(euler_flux(U))})
            rho2 = rho * float(np.exp(rng.uniform(0.05, 1.0) * rng.choice([-1, 1])))
            p2 = p * float(np.exp(rng.uniform(0.05, 1.0) * rng.choice([-1, 1])))
            vel2 = [float(v) for v in rng.uniform(-2, 2, size=Nd)]
            U2 = [float(u) for u in prim2cons(rho2, vel2, p2)]
            data["chandrashekar"].append({"Nd": Nd, "UL": Uf, "UR": U2,
                                          "F": f2(chandrashekar(Uf, U2))})
    OUT.write_text(json.dumps(data, indent=1))
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
