"""GPU tests of the mixed-precision batched path (p2de_batch_*) against the FP64 oracle.

The mixed path computes the same scheme with FP32 flux differencing, surface and low-order
fluxes and element assembly (FP64 node pressures, subcell flux potentials, limiter and RK
stages), so it is compared with tolerances that FP32 rounding explains:
  * rhsH and rhsL are sums of FP32 pair fluxes whose terms are as large as the physical
    flux times the operator row (|f| sum_j |S_ij| / (wq J)), while the sum can be far
    smaller (smooth flow, constant pressure): per component the worst node error is
    bounded by 32 FP32 unit roundoffs (2^-24) of that term scale;
  * the stage-1 CFL dt: relative 1e-6;
  * rhsU with the intended bounds (continuous in the data): as rhsH;
  * positivity (rho > 0, p > 0 in FP64 on the stored state) exactly after every stage;
  * conservation of every component (periodic, FP64 flux potentials telescoping exactly);
  * results do not depend on the batch composition, the launch shape or graph replay (bits).
"""
import numpy as np
import pytest

from conftest import PKG, normwise_rel, perturbed_state, problem, rel2
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

EPS32 = 2.0 ** -24
TOL_TERMS = 32 * EPS32


def term_scale(fp, U):
    """Per component: max_nodes |f_c| * max_i sum_j |S_ij| * metric / min(wq J), the size of the
    terms each node's RHS sums (Euler fluxes in both directions, LGL operators)."""
    g = fp.gamma
    rho, m1, m2, E = (U[..., c] for c in range(4))
    p = (g - 1) * (E - 0.5 * (m1 * m1 + m2 * m2) / rho)
    Fx = np.stack([m1, m1 * m1 / rho + p, m1 * m2 / rho, (E + p) * m1 / rho], -1)
    Fy = np.stack([m2, m1 * m2 / rho, m2 * m2 / rho + p, (E + p) * m2 / rho], -1)
    Fmax = np.maximum(np.abs(Fx).reshape(-1, 4).max(0), np.abs(Fy).reshape(-1, 4).max(0))
    Nh, K = fp.Nh, fp.K
    G = fp.GJh.reshape(-1, K, Nh)[:, :, 0]
    Sr = np.abs(fp.Srsh_db[:Nh * Nh].reshape(Nh, Nh).T).sum(1).max() * np.abs(G[0]).max()
    Ss = np.abs(fp.Srsh_db[Nh * Nh:].reshape(Nh, Nh).T).sum(1).max() * np.abs(G[3]).max()
    wJ = (fp.wq[None, :] * fp.Jq.reshape(K, fp.Nq)).min()
    return Fmax * max(Sr, Ss) / wJ


def term_err(a, b, scale):
    d = np.abs(np.asarray(a).reshape(-1, 4) - np.asarray(b).reshape(-1, 4)).max(0)
    return float(np.max(d / scale))

CASES = [
    ("vortex_2d", dict(K=8)),
    ("vortex_2d", dict(K=6, N=1)),
    ("vortex_2d", dict(K=6, N=2)),
    ("vortex_2d", dict(K=5, N=4)),
    ("riemann4_2d", dict(K=12)),
    ("kelvin_helmholtz", dict(K=8)),
    ("periodic_wave", dict(Kx=12, Ky=10, N=3, kx=2, ky=1, u0=0.3, v0=-0.7, dip=(0.4, 0.6), dt0=2e-5)),
]


def _ids(c):
    return f"{c[0]}-" + "-".join(f"{k}{v}" for k, v in c[1].items() if k in ("K", "N", "Kx"))


def mixed_fp(name, mode=0, **kw):
    solver, state, sp, fp = problem(name, **kw)
    fp.upper_bound_mode = mode
    return solver, state, sp, fp


@pytest.mark.parametrize("case", CASES, ids=[_ids(c) for c in CASES])
@pytest.mark.parametrize("mode", [0, 1], ids=["ref-bounds", "intended-bounds"])
def test_mixed_rhs_against_oracle(case, mode):
    from p2de_jl_amd.solver import BatchSolver, FIELD_RHSH, FIELD_RHSL
    name, kw = case
    solver, state, sp, fp = mixed_fp(name, mode, **kw)
    print(f"case {name} {kw} mode {mode}", flush=True)
    bs = BatchSolver([fp])
    bs.set_debug(True)
    orc = Oracle(fp)
    U0 = state.preallocation.Uq
    Nc = fp.Nc
    for trial, U in enumerate([U0, perturbed_state(U0, seed=3)]):
        bs.set_state(0, U)
        for nstage, dt in [(1, 0.9 * fp.CFL * fp.dt0), (2, 0.5 * fp.CFL * fp.dt0)]:
            dt_gpu, r_gpu = bs.rhs(0, 0.0, dt, nstage)
            dt_cpu = orc.rhs(U, 0.0, dt, nstage)
            sc = term_scale(fp, U)
            eH2 = rel2(bs.field(0, FIELD_RHSH), orc.field("rhsH"), Nc)
            eL2 = rel2(bs.field(0, FIELD_RHSL), orc.field("rhsL"), Nc)
            eU2 = rel2(r_gpu, orc.field("rhsU"), Nc)
            eH = term_err(bs.field(0, FIELD_RHSH), orc.field("rhsH"), sc)
            eL = term_err(bs.field(0, FIELD_RHSL), orc.field("rhsL"), sc)
            eU = term_err(r_gpu, orc.field("rhsU"), sc)
            edt = abs(dt_gpu - dt_cpu) / dt_cpu
            print(f"  trial {trial} stage {nstage}: relative 2-norm rhsH {eH2:.2e} rhsL {eL2:.2e} rhsU {eU2:.2e}; "
                  f"worst node / term scale: rhsH {eH:.2e} rhsL {eL:.2e} rhsU {eU:.2e}; dt {edt:.2e}", flush=True)
            assert max(eH, eL) <= TOL_TERMS
            assert edt <= 1e-6
            if mode == 1:
                assert eU <= TOL_TERMS


def _instances():
    """Three instances of different degree; the second has a near-vacuum dip."""
    out = []
    for kw in (dict(Kx=10, Ky=7, N=2, kx=1, ky=2, u0=0.5, v0=0.25),
               dict(Kx=9, Ky=11, N=3, kx=3, ky=1, u0=-0.8, v0=0.6, dip=(0.3, 0.7)),
               dict(Kx=6, Ky=8, N=4, kx=2, ky=2, u0=0.1, v0=-0.9)):
        param, ic, bc = PKG.configs.periodic_wave(**kw)
        solver, state, sp = PKG.initialize_DG(param, ic, bc)
        fp = PKG.flat_problem(solver, sp)
        fp.dt0 = 0.9 * PKG.timestep.cfl_dt(fp, state.preallocation.Uq) / fp.CFL
        out.append((fp, state.preallocation.Uq.copy()))
    return out


def _mass(fp, U):
    wJ = fp.wq[None, :] * fp.Jq.reshape(fp.K, fp.Nq)
    return np.einsum("kq,kqc->c", wJ, U)


def test_batch_positivity_conservation_and_independence():
    from p2de_jl_amd.solver import BatchSolver
    inst = _instances()
    fps = [fp for fp, _ in inst]
    bs = BatchSolver(fps)
    for i, (fp, U) in enumerate(inst):
        bs.set_state(i, U)
    bs.set_time(0.0)
    steps = 20
    bs.enqueue_steps(steps)
    bs.synchronize()
    res = [bs.get_state(i) for i in range(len(inst))]
    for i, (fp, U) in enumerate(inst):
        t, dt, n = bs.get_time(i)
        rep = bs.check_positivity(i)
        m0, m1 = _mass(fp, U), _mass(fp, res[i])
        drift = np.max(np.abs(m1 - m0) / np.maximum(np.abs(m0), 1e-300))
        print(f"  instance {i}: N={fp.N1D - 1} K={fp.K} t={t:.4e} steps={n} violations={rep.violations} "
              f"min rho {rep.min_rho:.3e} min p {rep.min_p:.3e} mass drift {drift:.2e}", flush=True)
        assert n == steps and rep.violations == 0 and rep.min_rho > 0 and rep.min_p > 0
        assert drift <= 1e-12
    # each instance alone gives the same bits
    for i, (fp, U) in enumerate(inst):
        one = BatchSolver([fp])
        one.set_state(0, U)
        one.enqueue_steps(steps)
        one.synchronize()
        assert np.array_equal(one.get_state(0), res[i]), f"instance {i} differs when run alone"
    # direct launches (no graph) and every launch shape give the same bits
    bs.set_graph(False)
    for v in range(bs.num_variants()):
        bs.set_variant(v)
        for i, (fp, U) in enumerate(inst):
            bs.set_state(i, U)
        bs.set_time(0.0)
        bs.enqueue_steps(steps)
        bs.synchronize()
        print(f"  variant {bs.variant()}", flush=True)
        for i in range(len(inst)):
            assert np.array_equal(bs.get_state(i), res[i])


def test_batch_steps_at_once_and_one_at_a_time_agree():
    """Steps enqueued all at once (one start launch, then the step graph replayed, each step's
    end starting the next) and one at a time give the same states and times bit for bit, and a
    state written between enqueues is the one whose time step is taken."""
    from p2de_jl_amd.solver import BatchSolver
    inst = _instances()
    fps = [fp for fp, _ in inst]
    steps = 6
    once, single = BatchSolver(fps), BatchSolver(fps)
    for bs in (once, single):
        for i, (fp, U) in enumerate(inst):
            bs.set_state(i, U)
        bs.set_time(0.0)
    once.enqueue_steps(steps)
    for k in range(steps):
        single.enqueue_steps(1)
    once.synchronize()
    single.synchronize()
    for i in range(len(inst)):
        ta, dta, na = once.get_time(i)
        tb, dtb, nb = single.get_time(i)
        print(f"  instance {i}: t {ta!r} vs {tb!r}, last dt {dta!r} vs {dtb!r}", flush=True)
        assert (ta, dta, na) == (tb, dtb, nb) and na == steps
        assert np.array_equal(once.get_state(i), single.get_state(i))
    # a new state between enqueues: its own wave speeds set the next dt
    fresh = BatchSolver(fps)
    for i, (fp, U) in enumerate(inst):
        U2 = U.copy()
        U2[..., 3] *= 1.5            # hotter gas: faster waves, a smaller dt
        once.set_state(i, U2)
        fresh.set_state(i, U2)
    t0 = [once.get_time(i)[0] for i in range(len(inst))]
    for i in range(len(inst)):
        fresh.set_time(t0[i], i)
    once.enqueue_steps(1)
    fresh.enqueue_steps(1)
    for i in range(len(inst)):
        t1, dt1, _ = once.get_time(i)
        t2, dt2, _ = fresh.get_time(i)
        assert dt1 == dt2 and t1 == t2
        assert np.array_equal(once.get_state(i), fresh.get_state(i))


def test_batch_autotune_cache_keeps_the_shape(tmp_path):
    """The autotuner's cache stores the launch shape (threads, elements per tile per degree):
    a second call reads it back; a shape this build does not have is timed again; lines of
    another kind (an FP64 context's shape index) are kept."""
    from p2de_jl_amd.solver import BatchSolver
    inst = _instances()
    bs = BatchSolver([fp for fp, _ in inst])
    for i, (fp, U) in enumerate(inst):
        bs.set_state(i, U)
    bs.set_time(0.0)
    cache = tmp_path / "tune.txt"
    cache.write_text("fp64:other-problem 3\n")
    v = bs.autotune(1, str(cache))
    lines = cache.read_text().splitlines()
    assert "fp64:other-problem 3" in lines
    (line,) = [l for l in lines if l.startswith("mixed")]
    key, shape = line.split()
    _, NT, E = bs.variant()
    assert shape == f"{NT}:{E[2]},{E[3]},{E[4]},{E[5]}"
    assert bs.autotune(1, str(cache)) == v
    assert bs.variant()[0] == v
    cache.write_text(f"{key} 4096:0,1,1,1\n")
    v2 = bs.autotune(1, str(cache))
    assert 0 <= v2 < bs.num_variants()
    (line2,) = [l for l in cache.read_text().splitlines() if l.startswith("mixed")]
    assert line2.split()[1] != "4096:0,1,1,1"
    # the state and time the tuner borrowed are restored
    for i, (fp, U) in enumerate(inst):
        assert np.array_equal(bs.get_state(i), U)
        assert bs.get_time(i)[0] == 0.0


def test_mixed_trajectory_against_fp64():
    """20 SSP-RK3 steps of the vortex on a 16x16 mesh: mixed against the FP64 drop-in."""
    from p2de_jl_amd.solver import BatchSolver, DeviceSolver
    solver, state, sp, fp = problem("vortex_2d", K=16)
    U0 = state.preallocation.Uq
    dev = DeviceSolver(fp)
    dev.set_state(U0)
    dev.set_time(0.0)
    dev.enqueue_steps(20)
    dev.synchronize()
    U64 = dev.get_state()
    bs = BatchSolver([fp])
    bs.set_state(0, U0)
    bs.set_time(0.0)
    bs.enqueue_steps(20)
    bs.synchronize()
    Um = bs.get_state(0)
    wJ = fp.wq[None, :] * fp.Jq.reshape(fp.K, fp.Nq)
    l1 = np.sum(wJ * np.abs(Um[..., 0] - U64[..., 0])) / np.sum(wJ * np.abs(U64[..., 0]))
    linf = np.max(np.abs(Um[..., 0] - U64[..., 0])) / np.max(np.abs(U64[..., 0]))
    t64 = dev.get_time()[0]
    tm = bs.get_time(0)[0]
    print(f"  density L1 {l1:.2e} Linf {linf:.2e}  t {tm:.10e} vs {t64:.10e}", flush=True)
    assert abs(tm - t64) <= 1e-6 * t64
    assert l1 <= 1e-6 and linf <= 1e-5


def test_single_context_mixed_precision_is_the_one_instance_batch():
    """p2de_create with precision MIXED forwards to a one-instance batch: the same bits as
    BatchSolver for rhs, fields, steps, time, run statistics and the positivity report."""
    from p2de_jl_amd.solver import BatchSolver, DeviceSolver, FIELD_RHSH, FIELD_RHSL, PRECISION_MIXED
    solver, state, sp, fp = problem("riemann4_2d", K=10)
    U0 = state.preallocation.Uq
    dev = DeviceSolver(fp, precision=PRECISION_MIXED)
    bs = BatchSolver([fp])
    dev.set_debug(True)
    bs.set_debug(True)
    dev.set_state(U0)
    bs.set_state(0, U0)
    dt_d, r_d = dev.rhs(0.0, 0.9 * fp.CFL * fp.dt0, 1)
    dt_b, r_b = bs.rhs(0, 0.0, 0.9 * fp.CFL * fp.dt0, 1)
    assert dt_d == dt_b and np.array_equal(r_d, r_b)
    assert np.array_equal(dev.field(FIELD_RHSH).ravel(), bs.field(0, FIELD_RHSH).ravel())
    assert np.array_equal(dev.field(FIELD_RHSL).ravel(), bs.field(0, FIELD_RHSL).ravel())
    t, dt = dev.step(0.0)
    bs.set_time(0.0)
    bs.enqueue_steps(1)
    bs.synchronize()
    tb, dtb, nb = bs.get_time(0)
    print(f"  step: t {t:.6e} dt {dt:.6e} vs batch t {tb:.6e} dt {dtb:.6e}", flush=True)
    assert t == tb and dt == dtb and nb == 1
    assert np.array_equal(dev.get_state(), bs.get_state(0))
    stats = dev.run(max_steps=7)
    bs.enqueue_steps(7)
    bs.synchronize()
    assert stats.steps == 7 and stats.t == bs.get_time(0)[0]
    assert np.array_equal(dev.get_state(), bs.get_state(0))
    rd, rb = dev.check_positivity(), bs.check_positivity(0)
    assert rd.violations == rb.violations == 0 and rd.min_rho == rb.min_rho and rd.min_p == rb.min_p
