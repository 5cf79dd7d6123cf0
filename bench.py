"""Benchmark: DOF-stage updates per second of the limited-ESDG SSP-RK3 step on one MI355X.

Headline workload (``value``): the batched workload -- 48 independent 2D periodic Euler solves
(element grids Kx, Ky in {32..160}, N in {2,3,4}, seed 0; about 7.4M DOFs) advanced together
in mixed precision (FP32 bulk, FP64 limiter, RK stages and state), one fused launch per RK
stage for all instances, steps replayed from a hipGraph, launch shape autotuned.

Beside it, on the 2D N=3 quad configurations (vortex 64x64, four-state Riemann 256x256), the
FP64 drop-in (``p2de_create``) and the mixed path are timed on the same steps, with their
ratio and the density L1/Linf difference of the two results (the batched instances are also
run one by one through the FP64 drop-in for that comparison).  The Gauss basis with the
entropy-projection limiter is timed through the FP64 drop-in on the four-state Riemann problem
(256x256, N=3) and on the reference's Kelvin-Helmholtz example (80x80, N=3), and the other limiter
set-ups (Zhang-Shu, shock capture, min-entropy bounds) on the four-state Riemann problem.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-compare]
    python bench.py --profile WORKLOAD [--precision mixed|fp64] [--basis B] [--variant V] [--steps K]
        (runs K steps of one workload and nothing else: the command the rocprofv3
         summaries in profiles/ were taken with)

Prints one JSON line.  ``value`` = DOFs x 3 stages x K steps / wall seconds of the batch.
``roofline`` is for the dominant kernel (the fused stage kernel): algorithmic bytes = 96 B per
DOF-stage (a step reads U^n three times and the two intermediate stages once each and writes
three stages: 9 FP64 state vectors of 32 B per 3 stages) x the DOFs one launch covers, over its
average duration from HIP events on the library's stream.  ``traffic`` is the HBM traffic per
launch measured with rocprofv3 (FETCH_SIZE, WRITE_SIZE, calibrated on a device copy) for the
same workload, read from profiles/ when recorded there.  ``cpu_baseline`` times the CPU oracle
(a restatement of the reference loops, OpenMP over elements) on the host on a bounded sample.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "DOF-stage updates/s, 2D Euler N=3: mixed/batched vs FP64 at matched L1 error"
BYTES_PER_DOF_STAGE = 96
PROFILE = ROOT / "profiles" / "mix_stage_batched48_traffic.md"


def profiled_traffic():
    """HBM bytes per fused-stage launch of the batched workload, as recorded by the rocprofv3
    counter passes summarised in profiles/ (None if not recorded)."""
    if not PROFILE.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", PROFILE.read_text())
    return float(m.group(1)) if m else None


def density_diff(fp, Um, U64):
    """Relative L1 (quadrature-weighted) and Linf difference of the densities."""
    import numpy as np
    wJ = fp.wq[None, :] * fp.Jq.reshape(fp.K, fp.Nq)
    d = Um[..., 0] - U64[..., 0]
    l1 = float(np.sum(wJ * np.abs(d)) / np.sum(wJ * np.abs(U64[..., 0])))
    linf = float(np.max(np.abs(d)) / np.max(np.abs(U64[..., 0])))
    return l1, linf


def time_batch(BatchSolver, fps, Us, steps, warmup, tune=True):
    bs = BatchSolver(fps)
    for i, U in enumerate(Us):
        bs.set_state(i, U)
    bs.set_time(0.0)
    if tune:
        bs.autotune(3)
    bs.enqueue_steps(warmup)
    bs.synchronize()
    for i, U in enumerate(Us):
        bs.set_state(i, U)
    bs.set_time(0.0)
    bs.synchronize()
    t0 = time.perf_counter()
    bs.enqueue_steps(steps)
    bs.synchronize()
    wall = time.perf_counter() - t0
    return bs, wall


def time_fp64(DeviceSolver, fp, U, steps, warmup):
    dev = DeviceSolver(fp)
    dev.set_state(U)
    dev.set_time(0.0)
    dev.autotune(3)
    dev.enqueue_steps(warmup)
    dev.synchronize()
    dev.set_state(U)
    dev.set_time(0.0)
    dev.synchronize()
    t0 = time.perf_counter()
    dev.enqueue_steps(steps)
    dev.synchronize()
    return dev, time.perf_counter() - t0


def cpu_baseline(pkg, fp, U, min_seconds=10.0):
    """Time the oracle on the host cores: one warm-up step, then steps until min_seconds pass."""
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    O.set_threads(threads)
    orc = O.Oracle(fp)
    U = U.copy()
    t, _ = orc.step(U, 0.0)
    steps = 0
    t0 = time.perf_counter()
    while True:
        t, _ = orc.step(U, t)
        steps += 1
        el = time.perf_counter() - t0
        if el >= min_seconds or steps >= 1000:
            break
    dofs = fp.K * fp.Nq
    return {"value": dofs * 3 * steps / el, "unit": "DOF-stage updates/s", "cores": O.get_threads(),
            "kind": "port",
            "sample": f"batched instance 0 ({fp.Kx}x{fp.Ky} elements, N={fp.N1D - 1}, {dofs} DOFs), FP64: "
                      f"{steps} SSP-RK3 steps after 1 warm-up step, {el:.2f} s wall on {O.get_threads()} "
                      f"OpenMP threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 comparison legs")
    ap.add_argument("--profile", choices=["batched", "vortex_2d", "riemann4_2d", "kelvin_helmholtz",
                                          "kelvin_helmholtz_gauss"],
                    help="only run --steps steps of this workload (for rocprofv3) and exit")
    ap.add_argument("--basis", choices=["lobatto", "gauss", "gauss_noep"], default=None,
                    help="approximation basis of a --profile workload (default: the workload's own)")
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--variant", type=int, default=0, help="launch shape for --profile")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB mebibytes (calibrates rocprof "
                         "FETCH_SIZE/WRITE_SIZE against a known byte count) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU decomposition is documented, not run.", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import torch
        n = args.copy_calibration << 20
        src = torch.empty(n, dtype=torch.uint8, device="cuda")
        src.fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n, "copies": 3}), flush=True)
        return

    import __graft_entry__ as G
    pkg = G.load_package()
    from p2de_jl_amd.solver import BatchSolver, DeviceSolver

    if args.profile:
        if args.profile == "batched":
            inst = pkg.configs.batched_instances()
        else:
            kw = {"basis": args.basis} if args.basis else {}
            inst = [pkg.configs.build(args.profile, **kw)]
        fps = [pkg.flat_problem(s, sp) for s, st, sp in inst]
        Us = [st.preallocation.Uq for s, st, sp in inst]
        if args.precision == "fp64":
            if len(fps) != 1:
                sys.exit("bench.py: the FP64 drop-in runs one instance")
            dev = DeviceSolver(fps[0])
            dev.set_variant(args.variant)
            dev.set_state(Us[0])
            dev.set_time(0.0)
            dev.enqueue_steps(args.steps)
            dev.synchronize()
            desc = dev.variant()
        else:
            dev = BatchSolver(fps)
            dev.set_variant(args.variant)
            for i, U in enumerate(Us):
                dev.set_state(i, U)
            dev.set_time(0.0)
            dev.enqueue_steps(args.steps)
            dev.synchronize()
            desc = dev.variant()
        print(json.dumps({"profile": args.profile, "basis": args.basis, "precision": args.precision, "variant": desc,
                          "steps": args.steps, "dofs": sum(fp.K * fp.Nq for fp in fps)}), flush=True)
        return

    # ---------------------------------------------------------------- headline: the batch
    inst = pkg.configs.batched_instances()
    fps = [pkg.flat_problem(s, sp) for s, st, sp in inst]
    Us = [st.preallocation.Uq for s, st, sp in inst]
    bs, wall = time_batch(BatchSolver, fps, Us, args.steps, args.warmup)
    dofs = bs.num_dofs(-1)
    value = dofs * 3 * args.steps / wall
    Um = [bs.get_state(i) for i in range(len(fps))]
    viol = 0
    min_rho = min_p = float("inf")
    for i in range(len(fps)):
        rep = bs.check_positivity(i)
        viol += int(rep.violations)
        min_rho = min(min_rho, rep.min_rho)
        min_p = min(min_p, rep.min_p)
    v, NT, E = bs.variant()
    # per-kernel durations (HIP events on the library's stream), separate pass
    for i, U in enumerate(Us):
        bs.set_state(i, U)
    bs.set_time(0.0)
    bs.set_kernel_timing(True)
    bs.enqueue_steps(min(args.steps, 10))
    ms, cnt = bs.kernel_timing()
    bs.set_kernel_timing(False)
    avg_stage = ms[0] / max(cnt[0], 1) * 1e-3
    avg_cfl = ms[1] / max(cnt[1], 1) * 1e-3
    alg_bytes = BYTES_PER_DOF_STAGE * dofs
    achieved = alg_bytes / avg_stage / 1e9
    traffic = profiled_traffic()

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 bulk + f64 limiter/RK/state",
        "data": "synthetic (analytic initial conditions of the named problems, seed 0)",
        "config": {"workload": "batched: 48 periodic 2D Euler solves, Kx,Ky in {32..160}, N in {2,3,4}, "
                               "seed 0, mixed precision, hipGraph RK3", "instances": len(fps), "dofs": dofs,
                   "dofs_by_degree": {f"N{n}": sum(fp.K * fp.Nq for fp in fps if fp.N1D == n + 1) for n in (2, 3, 4)},
                   "limiter": "SubcellLimiter(PositivityBound), reference bounds",
                   "launch_shape": {"variant": v, "threads": NT, "elements_per_tile": {f"N{n - 1}": E[n] for n in (3, 4, 5)}},
                   "positivity_violations": viol, "min_rho": min_rho, "min_p": min_p},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "mix_stage_kernel (fused RHS + limiter + RK stage)",
                     "avg_launch_us": avg_stage * 1e6, "algorithmic_bytes_per_launch": alg_bytes,
                     "cfl_pass_avg_us": avg_cfl * 1e6},
    }
    del bs

    if not args.no_compare:
        cmp = {}
        # the batch through the FP64 drop-in, instance by instance, for the density difference
        l1s, linfs, fp64_wall = [], [], 0.0
        for i, (fp, U) in enumerate(zip(fps, Us)):
            dev = DeviceSolver(fp)
            dev.set_state(U)
            dev.set_time(0.0)
            dev.synchronize()
            t0 = time.perf_counter()
            dev.enqueue_steps(args.steps)
            dev.synchronize()
            fp64_wall += time.perf_counter() - t0
            l1, linf = density_diff(fp, Um[i], dev.get_state())
            l1s.append(l1)
            linfs.append(linf)
            del dev
        cmp["batched"] = {"fp64_dropin_one_by_one": dofs * 3 * args.steps / fp64_wall, "mixed_batched": value,
                          "ratio": value / (dofs * 3 * args.steps / fp64_wall),
                          "density_L1_max_over_instances": max(l1s), "density_Linf_max_over_instances": max(linfs)}
        # the same batch with the bounds the bound tuple means (upper density bound Inf): the
        # limiter's per-face minimum rarely short-circuits, and its energy quadratic is solved
        fpi = [dataclasses.replace(fp, upper_bound_mode=1) for fp in fps]
        bi, wi = time_batch(BatchSolver, fpi, Us, args.steps, args.warmup)
        vi = sum(bi.check_positivity(i).violations for i in range(len(fpi)))
        cmp["batched_intended_bounds"] = {"mixed_batched": dofs * 3 * args.steps / wi, "ms_per_step": wi / args.steps * 1e3,
                                          "launch_variant": bi.variant()[0], "positivity_violations": int(vi)}
        del bi
        for name in ("vortex_2d", "riemann4_2d"):
            solver, state, sp = pkg.configs.build(name)
            fp = pkg.flat_problem(solver, sp)
            U = state.preallocation.Uq
            n = fp.K * fp.Nq
            dev, w64 = time_fp64(DeviceSolver, fp, U, args.steps, args.warmup)
            U64 = dev.get_state()
            rep64 = dev.check_positivity()
            del dev
            b1, wm = time_batch(BatchSolver, [fp], [U], args.steps, args.warmup)
            l1, linf = density_diff(fp, b1.get_state(0), U64)
            repm = b1.check_positivity(0)
            del b1
            r64 = n * 3 * args.steps / w64
            rm = n * 3 * args.steps / wm
            cmp[name] = {"dofs": n, "fp64_dropin": r64, "mixed": rm, "ratio": rm / r64, "density_L1": l1,
                         "density_Linf": linf, "min_p_fp64": rep64.min_p, "min_p_mixed": repm.min_p,
                         "positivity_violations": int(rep64.violations) + int(repm.violations)}
        # the Gauss basis (GaussCollocation + NodewiseScaledExtrapolation), FP64 drop-in: the
        # four-state Riemann problem at 256x256 N=3 and the reference's Kelvin-Helmholtz example
        gauss = {}
        for name, kw in (("riemann4_2d", {"basis": "gauss"}), ("kelvin_helmholtz_gauss", {})):
            solver, state, sp = pkg.configs.build(name, **kw)
            fp = pkg.flat_problem(solver, sp)
            U = state.preallocation.Uq
            n = fp.K * fp.Nq
            dev, w64 = time_fp64(DeviceSolver, fp, U, args.steps, args.warmup)
            rep = dev.check_positivity()
            dev.set_state(U)
            dev.set_time(0.0)
            dev.set_kernel_timing(True)
            dev.enqueue_steps(min(args.steps, 10))
            dev.synchronize()
            ms, cnt = dev.kernel_timing()
            dev.set_kernel_timing(False)
            steps_t = max(cnt[1], 1) / 3
            del dev
            gauss[name] = {"dofs": n, "fp64_dropin": n * 3 * args.steps / w64,
                           "ms_per_step": w64 / args.steps * 1e3,
                           "projection_and_element_kernels_ms_per_step": ms[0] / steps_t,
                           "update_kernel_ms_per_step": ms[1] / steps_t,
                           "positivity_violations": int(rep.violations), "min_p": rep.min_p}
        cmp["gauss_fp64"] = gauss
        # the other limiter set-ups of test/test_smoke.jl:44-52 and the TVD bounds, FP64 drop-in,
        # four-state Riemann 256x256 N=3
        lims = {}
        for lim in ("zhangshu", "zhangshu_shock", "subcell_shock", "minentropy", "relaxed_minentropy",
                    "cellentropy", "relaxed_cellentropy", "tvd", "tvd_minentropy"):
            solver, state, sp = pkg.configs.build("riemann4_2d", limiter=lim)
            fp = pkg.flat_problem(solver, sp)
            U = state.preallocation.Uq
            n = fp.K * fp.Nq
            dev, w64 = time_fp64(DeviceSolver, fp, U, args.steps, args.warmup)
            rep = dev.check_positivity()
            del dev
            lims[lim] = {"fp64_dropin": n * 3 * args.steps / w64, "ms_per_step": w64 / args.steps * 1e3,
                         "positivity_violations": int(rep.violations)}
        cmp["limiters_fp64_riemann4"] = lims
        out["config"]["comparison"] = cmp
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(pkg, fps[0], Us[0])
        except Exception as e:  # the baseline is reported, never the thing measured
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
