"""The C-ABI library on the host (CPU, no kernel launches): it loads, exports every
symbol include/p2de_hip.h declares, and rejects bad arguments and unsupported options
with a status and a message instead of crashing (the checks run before any device call)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from conftest import PKG, problem

S = PKG.solver
HEADER = Path(__file__).resolve().parents[1] / "include" / "p2de_hip.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(p2de_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    L = S.lib()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(S.EXPORTED_SYMBOLS)


def test_status_strings():
    L = S.lib()
    assert L.p2de_status_string(0) == b"ok"
    assert b"unsupported" in L.p2de_status_string(S.P2DE_E_UNSUPPORTED)


def _create(fp, **cfg_over):
    """Call p2de_create like DeviceSolver does, with config overrides; return (status, msg)."""
    cfg = S._Config(Nd=fp.Nd, N=fp.N1D - 1, Kx=fp.Kx, Ky=fp.Ky, gamma=fp.gamma, zeta=fp.zeta, eta=fp.eta,
                    POSTOL=fp.POSTOL, ZEROTOL=fp.ZEROTOL, CFL=fp.CFL, dt0=fp.dt0, t0=fp.t0, T=fp.T,
                    high_order_surface_flux=1, basis=fp.basis, entropyproj_limiter=fp.entropyproj_limiter,
                    rhs_limiter=fp.rhs_limiter, bound=fp.bound, shockcapture=fp.shockcapture,
                    shock_a=fp.shock_a, shock_c=fp.shock_c, low_order_surface_flux=fp.low_order_surface_flux,
                    cell_beta=fp.cell_beta)
    for k, v in cfg_over.items():
        setattr(cfg, k, v)
    keep = {n: None if getattr(fp, n) is None else np.ascontiguousarray(getattr(fp, n), dtype=np.float64)
            for n in ("Srsh_db", "Srs0", "Brs", "MinvVhT", "MinvVfT", "Vf", "Vf_low", "wq", "Jq", "GJh", "Ival",
                      "VDM_inv")}
    keep.update({n: np.ascontiguousarray(getattr(fp, n), dtype=np.int64)
                 for n in ("Srsh_nnz", "Srs0_nnz", "fq2q", "q2fq", "mapP", "mapI", "mapO")})
    return _create_raw(fp, cfg, keep)


def _create_raw(fp, cfg, keep):
    d, i = S._d, S._i
    ops = S._Operators(Nq=fp.Nq, Nfp=fp.Nfp, Nh=fp.Nh, Srsh_db=d(keep["Srsh_db"]), Srs0=d(keep["Srs0"]),
                       nnz_h=len(keep["Srsh_nnz"]) // 2, Srsh_nnz=i(keep["Srsh_nnz"]),
                       nnz_0=len(keep["Srs0_nnz"]) // 2, Srs0_nnz=i(keep["Srs0_nnz"]), Brs=d(keep["Brs"]),
                       MinvVhT=d(keep["MinvVhT"]), MinvVfT=d(keep["MinvVfT"]), Vf=d(keep["Vf"]),
                       Vf_low=d(keep["Vf_low"]), wq=d(keep["wq"]), fq2q=i(keep["fq2q"]), q2fq=i(keep["q2fq"]),
                       VDM_inv=d(keep["VDM_inv"]))
    geo = S._Geometry(Jq=d(keep["Jq"]), GJh=d(keep["GJh"]))
    bc = S._BCData(mapP=i(keep["mapP"]), nI=len(keep["mapI"]), mapI=i(keep["mapI"]), Ival=d(keep["Ival"]),
                   nO=len(keep["mapO"]), mapO=i(keep["mapO"]))
    h = C.c_void_p()
    st = S.lib().p2de_create(C.byref(cfg), C.byref(ops), C.byref(geo), C.byref(bc), C.byref(h))
    msg = S.lib().p2de_last_error(None).decode()
    if st == 0:
        S.lib().p2de_destroy(h)
    return st, msg


@pytest.fixture(scope="module")
def fp2():
    return problem("vortex_2d", K=4)[3]


@pytest.mark.parametrize("field,value,code", [
    ("basis", 1, S.P2DE_E_UNSUPPORTED), ("rhs_limiter", 5, S.P2DE_E_UNSUPPORTED),
    ("entropyproj_limiter", 1, S.P2DE_E_UNSUPPORTED), ("shockcapture", 3, S.P2DE_E_UNSUPPORTED),
    ("bound", 10, S.P2DE_E_UNSUPPORTED),
    ("high_order_surface_flux", 0, S.P2DE_E_UNSUPPORTED), ("precision", 7, S.P2DE_E_INVALID),
    ("upper_bound_mode", 5, S.P2DE_E_INVALID), ("Nd", 3, S.P2DE_E_INVALID), ("N", 9, S.P2DE_E_UNSUPPORTED),
    ("Kx", 0, S.P2DE_E_INVALID), ("gamma", 1.0, S.P2DE_E_INVALID)])
def test_create_rejects_options(fp2, field, value, code):
    st, msg = _create(fp2, **{field: value})
    assert st == code and msg


def test_create_rejects_bad_operators_and_mesh(fp2):
    cfg_fields = dict(Nd=fp2.Nd, N=fp2.N1D - 1, Kx=fp2.Kx, Ky=fp2.Ky, gamma=fp2.gamma, zeta=fp2.zeta, eta=fp2.eta,
                      POSTOL=fp2.POSTOL, ZEROTOL=fp2.ZEROTOL, CFL=fp2.CFL, dt0=fp2.dt0, t0=fp2.t0, T=fp2.T,
                      high_order_surface_flux=1)

    def base():
        keep = {n: np.ascontiguousarray(getattr(fp2, n), dtype=np.float64).copy()
                for n in ("Srsh_db", "Srs0", "Brs", "MinvVhT", "MinvVfT", "Vf", "Vf_low", "wq", "Jq", "GJh", "Ival",
                          "VDM_inv")}
        keep.update({n: np.ascontiguousarray(getattr(fp2, n), dtype=np.int64).copy()
                     for n in ("Srsh_nnz", "Srs0_nnz", "fq2q", "q2fq", "mapP", "mapI", "mapO")})
        return keep

    cases = []
    k = base(); k["mapP"][3] = 10 ** 9; cases.append((k, S.P2DE_E_INVALID, "mapP"))
    k = base(); k["fq2q"][0] = 99; cases.append((k, S.P2DE_E_INVALID, "fq2q"))
    k = base(); k["Jq"][1] *= 2.0; cases.append((k, S.P2DE_E_UNSUPPORTED, "Jq"))
    k = base(); k["Srsh_nnz"] = k["Srsh_nnz"][:-2]; cases.append((k, S.P2DE_E_UNSUPPORTED, "Srsh_nnz"))
    k = base(); k["Vf"][fp2.Nfp * 3 + 1] = 0.5; cases.append((k, S.P2DE_E_UNSUPPORTED, "Vf"))
    k = base(); k["MinvVhT"][1] = 0.25; cases.append((k, S.P2DE_E_UNSUPPORTED, "MinvVhT"))
    K, Nh = fp2.K, fp2.Nh
    k = base(); k["GJh"][1 * Nh * K] = 0.1; cases.append((k, S.P2DE_E_UNSUPPORTED, "metric"))
    for keep, code, what in cases:
        st, msg = _create_raw(fp2, S._Config(**cfg_fields), keep)
        assert st == code, (what, st, msg)
        assert msg


def test_null_arguments():
    L = S.lib()
    h = C.c_void_p()
    assert L.p2de_create(None, None, None, None, C.byref(h)) == S.P2DE_E_INVALID
    assert L.p2de_set_state(None, None) == S.P2DE_E_INVALID
    assert L.p2de_rhs(None, 0.0, 0.0, 1, None, None) == S.P2DE_E_INVALID
    assert L.p2de_run(None, 1, None) == S.P2DE_E_INVALID
    assert L.p2de_num_dofs(None) == 0
    assert L.p2de_run_segment(None, 1, 0, None, None, 1, None, None) == S.P2DE_E_INVALID
    assert L.p2de_check_conservation(None, None) == S.P2DE_E_INVALID
    assert L.p2de_checkpoint_size(None) == 0
    assert L.p2de_checkpoint_save(None, 1, None, 0) == S.P2DE_E_INVALID
    assert L.p2de_checkpoint_load(None, None, 0, None) == S.P2DE_E_INVALID


# ----------------------------------------------------------------- batched C-ABI, host checks
def _batch_create(fps, **cfg_over):
    """p2de_batch_create on descriptors as BatchSolver builds them, with config overrides;
    returns (status, message).  The checks run before any device call."""
    n = len(fps)
    cfgs, opss, geos, bcs = (S._Config * n)(), (S._Operators * n)(), (S._Geometry * n)(), (S._BCData * n)()
    keep = []
    for j, fp in enumerate(fps):
        cfg, ops, geo, bc, kp = S._descriptors(fp, None, S.PRECISION_MIXED)
        for k, v in cfg_over.items():
            setattr(cfg, k, v)
        cfgs[j], opss[j], geos[j], bcs[j] = cfg, ops, geo, bc
        keep.append(kp)
    h = C.c_void_p()
    st = S.lib().p2de_batch_create(n, cfgs, opss, geos, bcs, C.byref(h))
    msg = S.lib().p2de_batch_last_error(None).decode()
    if st == 0:
        S.lib().p2de_batch_destroy(h)
    return st, msg


def test_batch_create_rejects_before_device_work():
    _, _, _, fp2 = problem("vortex_2d", K=4)
    _, _, _, fp1 = problem("sod_1d", K=8)
    L = S.lib()
    h = C.c_void_p()
    assert L.p2de_batch_create(0, None, None, None, None, C.byref(h)) == S.P2DE_E_INVALID
    st, msg = _batch_create([fp2], precision=S.PRECISION_FP64)
    assert st == S.P2DE_E_UNSUPPORTED and "mixed" in msg
    st, msg = _batch_create([fp1])
    assert st == S.P2DE_E_UNSUPPORTED and msg
    st, msg = _batch_create([fp2], N=9)
    assert st == S.P2DE_E_UNSUPPORTED and msg
    st, msg = _batch_create([fp2], upper_bound_mode=3)
    assert st == S.P2DE_E_INVALID and msg


def test_create_mixed_precision_goes_through_the_batch_checks():
    """p2de_create with precision MIXED runs a one-instance batch: the batch's own checks
    reject what the mixed path does not cover, with their message, before device work."""
    _, _, _, fp2 = problem("vortex_2d", K=4)
    _, _, _, fp1 = problem("sod_1d", K=8)
    st, msg = _create(fp1, precision=S.PRECISION_MIXED)
    assert st == S.P2DE_E_UNSUPPORTED and "2D" in msg
    st, msg = _create(fp2, precision=S.PRECISION_MIXED, N=9)
    assert st == S.P2DE_E_UNSUPPORTED and msg
    st, msg = _create(fp2, precision=S.PRECISION_MIXED, upper_bound_mode=3)
    assert st == S.P2DE_E_INVALID and msg


def test_gauss_options_and_operators_are_checked():
    """Gauss operators need basis = Gauss (their Vf rows interpolate); the Lobatto basis takes
    no entropy-projection limiter; the mixed-precision path runs Lobatto only."""
    _, _, _, fpg = problem("vortex_2d", K=4, basis="gauss")
    _, _, _, fpl = problem("vortex_2d", K=4)
    st, msg = _create(fpg, basis=S.BASIS_LOBATTO, entropyproj_limiter=0)
    assert st == S.P2DE_E_UNSUPPORTED and "Lobatto operator" in msg
    st, msg = _create(fpl, entropyproj_limiter=1)
    assert st == S.P2DE_E_UNSUPPORTED and "Lobatto" in msg
    st, msg = _create(fpg, precision=S.PRECISION_MIXED)
    assert st == S.P2DE_E_UNSUPPORTED and "Lobatto" in msg
    st, msg = _create(fpg, basis=7)
    assert st == S.P2DE_E_UNSUPPORTED and msg


def test_limiter_options_are_checked():
    """Zhang-Shu takes no subcell bound; the shock indicator and the relaxed bound need VDM_inv;
    the mixed-precision path runs SubcellLimiter(PositivityBound, NoShockCapture) only."""
    _, _, _, zs = problem("smoke_vortex", limiter="zhangshu_shock")
    st, msg = _create(zs, bound=1)
    assert st == S.P2DE_E_UNSUPPORTED and "Zhang" in msg
    _, _, _, rel = problem("smoke_vortex", limiter="relaxed_minentropy")
    rel.VDM_inv = None
    st, msg = _create(rel)
    assert st == S.P2DE_E_INVALID and "VDM_inv" in msg
    _, _, _, me = problem("smoke_vortex", limiter="minentropy")
    st, msg = _create(me, precision=S.PRECISION_MIXED)
    assert st == S.P2DE_E_UNSUPPORTED and "mixed" in msg
    # the cell-entropy bounds: 2D Lobatto only (the reference cannot run them in 1D, and on Gauss
    # their interface step depends on the thread schedule); the relaxed one reads the indicator
    _, _, _, ce = problem("smoke_vortex", limiter="relaxed_cellentropy")
    _, _, _, g1 = problem("sod_1d", K=8)
    _, _, _, gg = problem("smoke_vortex", basis="gauss")
    for fp in (g1, gg):
        for bound in (3, 4, 8, 9):
            st, msg = _create(fp, bound=bound)
            assert st == S.P2DE_E_UNSUPPORTED and "cell-entropy" in msg
    ce.VDM_inv = None
    st, msg = _create(ce)
    assert st == S.P2DE_E_INVALID and "VDM_inv" in msg
