"""Record dt0 = 0.9 * dt_pos(U0) / CFL for the benchmark problems.

dt_pos is the low-order positivity time step (the CFL dt the oracle's stage-1 rhs
returns when the cap CFL*dt0 is lifted).  Run from the repository root:
    python tests/golden/record_dt0.py
and paste the printed dictionary into p2de.jl_amd/configs.py (RECORDED_DT0).
"""
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
import __graft_entry__ as G  # noqa: E402

pkg = G.load_package()
from oracle.oracle import Oracle  # noqa: E402

out = {}
for name in ("sod_1d", "vortex_2d", "riemann4_2d", "kelvin_helmholtz"):
    solver, state, sp = pkg.configs.build(name, dt0=1e30)
    fp = pkg.flat_problem(solver, sp)
    dt_pos = Oracle(fp).rhs(state.preallocation.Uq, 0.0, 1e30, 1)
    out[name] = 0.9 * dt_pos / fp.CFL
    print(name, dt_pos, out[name], flush=True)
print("RECORDED_DT0 =", repr(out))
