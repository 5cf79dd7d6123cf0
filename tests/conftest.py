"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on the CPU."""
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

import __graft_entry__ as G  # noqa: E402

PKG = G.load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs on the MI355X box)")


@pytest.fixture(scope="session")
def pkg():
    return PKG


def problem(name, **kw):
    """(solver, state, state_param, flat problem) of a named configuration."""
    solver, state, sp = PKG.configs.build(name, **kw)
    return solver, state, sp, PKG.flat_problem(solver, sp)


def normwise_rel(a, b, Nc):
    """max over components of ||a_c - b_c||_inf / ||b_c||_inf (the parity metric)."""
    a = np.asarray(a).reshape(-1, Nc)
    b = np.asarray(b).reshape(-1, Nc)
    den = np.max(np.abs(b), axis=0)
    den = np.where(den == 0, 1.0, den)
    return float(np.max(np.max(np.abs(a - b), axis=0) / den))


def perturbed_state(U, seed=0, amp=0.05, gamma=1.4):
    """Seeded perturbation in primitive variables: rho and p scaled by 1 +- amp, velocities
    shifted by +- amp times the largest speed; rho and p stay positive for amp < 1."""
    rng = np.random.default_rng(seed)
    Nd = U.shape[-1] - 2
    rho = U[..., 0]
    vel = U[..., 1:1 + Nd] / rho[..., None]
    p = (gamma - 1) * (U[..., -1] - 0.5 * rho * np.sum(vel * vel, axis=-1))
    vmax = max(float(np.max(np.abs(vel))), 1e-3)
    rho = rho * (1 + amp * rng.uniform(-1, 1, size=rho.shape))
    p = p * (1 + amp * rng.uniform(-1, 1, size=p.shape))
    vel = vel + amp * vmax * rng.uniform(-1, 1, size=vel.shape)
    out = np.empty_like(U)
    out[..., 0] = rho
    out[..., 1:1 + Nd] = rho[..., None] * vel
    out[..., -1] = p / (gamma - 1) + 0.5 * rho * np.sum(vel * vel, axis=-1)
    return out


def rel2(a, b, Nc):
    """max over components of ||a_c - b_c||_2 / ||b_c||_2."""
    a = np.asarray(a).reshape(-1, Nc)
    b = np.asarray(b).reshape(-1, Nc)
    den = np.sqrt(np.sum(b * b, axis=0))
    den = np.where(den == 0, 1.0, den)
    return float(np.max(np.sqrt(np.sum((a - b) ** 2, axis=0)) / den))
