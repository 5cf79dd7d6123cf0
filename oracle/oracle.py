"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE).

Only the tests, ``__graft_entry__.smoke()`` (as the checker) and ``bench.py``'s
``cpu_baseline`` leg use this module.  It loads ``oracle/libp2de_oracle.so`` (built by
``oracle/Makefile``) and drives it with the same flat problem arrays the product's
C-ABI receives.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libp2de_oracle.so"

FIELDS = {
    "rhsU": 0, "rhsH": 1, "rhsL": 2, "rhsxyH": 3, "rhsxyL": 4, "L_local": 5, "BF_H": 6, "BF_L": 7,
    "u_tilde": 8, "QF1": 9, "lambda": 10, "lambdaB": 11, "f_bar_H": 12, "f_bar_L": 13, "vq": 14,
    "psi_tilde": 15, "fstar_H": 16, "fstar_L": 17, "theta_local": 18, "smooth_indicator": 19,
    "blending_factor": 20, "smooth_factor": 21, "lbound_s_modified": 22, "L_zhangshu": 23,
    "s_modified_min": 24, "lbound_rho": 25, "ubound_rho": 26,
}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int64)


class _Problem(C.Structure):
    _fields_ = [
        ("Nd", C.c_int), ("Nc", C.c_int), ("N1D", C.c_int), ("Nq", C.c_int), ("Nfp", C.c_int),
        ("Nh", C.c_int), ("K", C.c_int),
        ("gamma", C.c_double), ("zeta", C.c_double), ("eta", C.c_double), ("POSTOL", C.c_double),
        ("ZEROTOL", C.c_double), ("CFL", C.c_double), ("dt0", C.c_double), ("t0", C.c_double), ("T", C.c_double),
        ("upper_bound_mode", C.c_int), ("entropyproj_limiter", C.c_int), ("rhs_limiter", C.c_int),
        ("bound", C.c_int), ("shockcapture", C.c_int), ("shock_a", C.c_double), ("shock_c", C.c_double),
        ("cell_beta", C.c_double),
        ("low_order_surface_flux", C.c_int),
        ("Srsh_db", _dp), ("Srs0", _dp), ("nnz_h", C.c_int), ("Srsh_nnz", _ip), ("nnz_0", C.c_int),
        ("Srs0_nnz", _ip), ("Brs", _dp), ("MinvVhT", _dp), ("MinvVfT", _dp), ("Vf", _dp), ("Vf_low", _dp),
        ("wq", _dp), ("fq2q", _ip), ("q2fq", _ip), ("VDM_inv", _dp), ("Jq", _dp), ("GJh", _dp), ("mapP", _ip),
        ("nI", C.c_int), ("mapI", _ip), ("Ival", _dp), ("nO", C.c_int), ("mapO", _ip),
    ]


def build_oracle(force: bool = False) -> Path:
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-C", str(HERE), "-s"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build_oracle()
        L = C.CDLL(str(LIB_PATH))
        L.oracle_create.argtypes = [C.POINTER(_Problem), C.POINTER(C.c_void_p)]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_field_size.argtypes = [C.c_void_p, C.c_int]
        L.oracle_field_size.restype = C.c_int64
        L.oracle_field.argtypes = [C.c_void_p, C.c_int]
        L.oracle_field.restype = _dp
        L.oracle_rhs.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, C.c_int, _dp]
        L.oracle_ssprk3_step.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.oracle_count_nonpositive.argtypes = [C.c_void_p, _dp, _ip]
        L.oracle_count_nonpositive.restype = C.c_int64
        L.oracle_check_conservation.argtypes = [C.c_void_p, _dp]
        L.oracle_check_conservation.restype = C.c_double
        for nm in ("oracle_pt_v_of_u", "oracle_pt_u_of_v", "oracle_pt_flux", "oracle_pt_fS"):
            getattr(L, nm).restype = None
        L.oracle_pt_v_of_u.argtypes = [C.c_int, C.c_double, _dp, _dp]
        L.oracle_pt_u_of_v.argtypes = [C.c_int, C.c_double, _dp, _dp]
        L.oracle_pt_flux.argtypes = [C.c_int, C.c_double, _dp, _dp]
        L.oracle_pt_fS.argtypes = [C.c_int, C.c_double, _dp, _dp, _dp]
        L.oracle_pt_logmean.argtypes = [C.c_double, C.c_double]
        L.oracle_pt_logmean.restype = C.c_double
        L.oracle_pt_wavespeed.argtypes = [C.c_int, C.c_double, _dp, _dp]
        L.oracle_pt_wavespeed.restype = C.c_double
        L.oracle_pt_limiting_param.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp]
        L.oracle_pt_limiting_param.restype = C.c_double
        L.oracle_pt_find_alpha.argtypes = [C.c_int, C.c_double, _dp, _dp]
        L.oracle_pt_find_alpha.restype = C.c_double
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_get_threads.restype = C.c_int
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None and a.size else None


def _i(a):
    return a.ctypes.data_as(_ip) if a is not None and a.size else None


class Oracle:
    """One oracle context for a ``FlatProblem`` (``p2de_jl_amd.dg.flat_problem``)."""

    def __init__(self, fp, upper_bound_mode=None):
        self.fp = fp
        self._keep = {}
        for name in ("Srsh_db", "Srs0", "Brs", "MinvVhT", "MinvVfT", "Vf", "Vf_low", "wq", "Jq", "GJh", "Ival"):
            self._keep[name] = np.ascontiguousarray(getattr(fp, name), dtype=np.float64)
        vinv = getattr(fp, "VDM_inv", None)
        self._keep["VDM_inv"] = None if vinv is None else np.ascontiguousarray(vinv, dtype=np.float64)
        for name in ("Srsh_nnz", "Srs0_nnz", "fq2q", "q2fq", "mapP", "mapI", "mapO"):
            self._keep[name] = np.ascontiguousarray(getattr(fp, name), dtype=np.int64)
        k = self._keep
        mode = fp.upper_bound_mode if upper_bound_mode is None else upper_bound_mode
        self.prob = _Problem(
            Nd=fp.Nd, Nc=fp.Nc, N1D=fp.N1D, Nq=fp.Nq, Nfp=fp.Nfp, Nh=fp.Nh, K=fp.K,
            gamma=fp.gamma, zeta=fp.zeta, eta=fp.eta, POSTOL=fp.POSTOL, ZEROTOL=fp.ZEROTOL,
            CFL=fp.CFL, dt0=fp.dt0, t0=fp.t0, T=fp.T, upper_bound_mode=int(mode),
            entropyproj_limiter=int(getattr(fp, "entropyproj_limiter", 0)),
            rhs_limiter=int(getattr(fp, "rhs_limiter", 0)), bound=int(getattr(fp, "bound", 0)),
            shockcapture=int(getattr(fp, "shockcapture", 0)), shock_a=float(getattr(fp, "shock_a", 0.5)),
            shock_c=float(getattr(fp, "shock_c", 1.8)), cell_beta=float(getattr(fp, "cell_beta", 0.5)), VDM_inv=_d(self._keep["VDM_inv"]),
            low_order_surface_flux=int(getattr(fp, "low_order_surface_flux", 0)),
            Srsh_db=_d(k["Srsh_db"]), Srs0=_d(k["Srs0"]), nnz_h=len(k["Srsh_nnz"]) // 2,
            Srsh_nnz=_i(k["Srsh_nnz"]), nnz_0=len(k["Srs0_nnz"]) // 2, Srs0_nnz=_i(k["Srs0_nnz"]),
            Brs=_d(k["Brs"]), MinvVhT=_d(k["MinvVhT"]), MinvVfT=_d(k["MinvVfT"]), Vf=_d(k["Vf"]),
            Vf_low=_d(k["Vf_low"]), wq=_d(k["wq"]), fq2q=_i(k["fq2q"]), q2fq=_i(k["q2fq"]), Jq=_d(k["Jq"]),
            GJh=_d(k["GJh"]), mapP=_i(k["mapP"]), nI=len(k["mapI"]), mapI=_i(k["mapI"]), Ival=_d(k["Ival"]),
            nO=len(k["mapO"]), mapO=_i(k["mapO"]))
        h = C.c_void_p()
        err = lib().oracle_create(C.byref(self.prob), C.byref(h))
        if err:
            raise RuntimeError(f"oracle_create failed ({err})")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.oracle_destroy(self.h)
            self.h = None

    def rhs(self, Uq: np.ndarray, t: float, dt: float, nstage: int) -> float:
        U = np.ascontiguousarray(Uq, dtype=np.float64)
        out = C.c_double(0.0)
        err = lib().oracle_rhs(self.h, _d(U), float(t), float(dt), int(nstage), C.byref(out))
        if err:
            raise RuntimeError(f"oracle_rhs failed ({err})")
        return out.value

    def field(self, name: str) -> np.ndarray:
        f = FIELDS[name]
        n = lib().oracle_field_size(self.h, f)
        ptr = lib().oracle_field(self.h, f)
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

    def step(self, Uq: np.ndarray, t: float):
        """One SSP33! step in place on ``Uq`` (C-contiguous float64); returns (t, dt)."""
        assert Uq.flags.c_contiguous and Uq.dtype == np.float64
        tt = C.c_double(t)
        dt = C.c_double(0.0)
        err = lib().oracle_ssprk3_step(self.h, _d(Uq), C.byref(tt), C.byref(dt))
        if err:
            raise RuntimeError(f"oracle_ssprk3_step failed ({err})")
        return tt.value, dt.value

    def count_nonpositive(self, Uq):
        U = np.ascontiguousarray(Uq, dtype=np.float64)
        first = C.c_int64(-1)
        n = lib().oracle_count_nonpositive(self.h, _d(U), C.byref(first))
        return int(n), int(first.value)

    def conservation(self, Uq) -> float:
        U = np.ascontiguousarray(Uq, dtype=np.float64)
        return lib().oracle_check_conservation(self.h, _d(U))


def set_threads(n: int) -> None:
    lib().oracle_set_threads(int(n))


def get_threads() -> int:
    return int(lib().oracle_get_threads())


# ----------------------------------------------------------------------------- pointwise
def _vec(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def pt_v_of_u(Nd, gamma, U):
    U = _vec(U)
    V = np.empty(Nd + 2)
    lib().oracle_pt_v_of_u(Nd, gamma, _d(U), _d(V))
    return V


def pt_u_of_v(Nd, gamma, V):
    V = _vec(V)
    U = np.empty(Nd + 2)
    lib().oracle_pt_u_of_v(Nd, gamma, _d(V), _d(U))
    return U


def pt_flux(Nd, gamma, U):
    U = _vec(U)
    F = np.empty(Nd * (Nd + 2))
    lib().oracle_pt_flux(Nd, gamma, _d(U), _d(F))
    return F.reshape(Nd, Nd + 2)


def pt_fS(Nd, gamma, UL, UR):
    UL, UR = _vec(UL), _vec(UR)
    F = np.empty(Nd * (Nd + 2))
    lib().oracle_pt_fS(Nd, gamma, _d(UL), _d(UR), _d(F))
    return F.reshape(Nd, Nd + 2)


def pt_logmean(aL, aR):
    return lib().oracle_pt_logmean(float(aL), float(aR))


def pt_wavespeed(Nd, gamma, U, n):
    U, n = _vec(U), _vec(n)
    return lib().oracle_pt_wavespeed(Nd, gamma, _d(U), _d(n))


def pt_limiting_param(Nd, zeta, ZEROTOL, mode, uL, P):
    uL, P = _vec(uL), _vec(P)
    return lib().oracle_pt_limiting_param(Nd, zeta, ZEROTOL, mode, _d(uL), _d(P))


def pt_find_alpha(Nd, POSTOL, ui, ut):
    ui, ut = _vec(ui), _vec(ut)
    return lib().oracle_pt_find_alpha(Nd, POSTOL, _d(ui), _d(ut))
