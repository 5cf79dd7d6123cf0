"""GPU drop-ins for ``rhs!`` and ``SSP33!`` over the C-ABI library ``libp2de_hip.so``.

``DeviceSolver`` owns one native context (``p2de_create``) built from a ``FlatProblem``.
``rhs(state, solver, state_param, time_param)`` and ``SSP33(state, solver, state_param)``
mirror src/dg/rhs/rhs.jl:5-13 and src/timestepping/SSPRK33.jl:1-62 with the same
arguments and meaning; they run on the GPU and raise ``P2DEError`` with the library's
message on any failure.  There is no CPU fallback: without the built library every call
fails.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from .dg import FlatProblem, State, StateParam, Solver, TimeParam, flat_problem

LIB_PATH = Path(__file__).resolve().parent / "csrc" / "libp2de_hip.so"

P2DE_OK, P2DE_E_INVALID, P2DE_E_UNSUPPORTED, P2DE_E_HIP, P2DE_E_ALLOC, P2DE_E_STATE, P2DE_E_POSITIVITY = range(7)
FIELD_RHSU, FIELD_RHSH, FIELD_RHSL, FIELD_L_LOCAL, FIELD_THETA, FIELD_L, FIELD_THETA_MEAN = range(7)
UPPER_BOUND_REFERENCE, UPPER_BOUND_INTENDED = 0, 1
PRECISION_FP64, PRECISION_MIXED = 0, 1
BASIS_LOBATTO, BASIS_GAUSS = 0, 1


class P2DEError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"[{status}] {msg}")
        self.status = status


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int64)


class _Config(C.Structure):
    _fields_ = [("Nd", C.c_int), ("N", C.c_int), ("Kx", C.c_int), ("Ky", C.c_int),
                ("gamma", C.c_double), ("zeta", C.c_double), ("eta", C.c_double), ("POSTOL", C.c_double),
                ("ZEROTOL", C.c_double), ("CFL", C.c_double), ("dt0", C.c_double), ("t0", C.c_double),
                ("T", C.c_double),
                ("rhs", C.c_int), ("low_order_surface_flux", C.c_int), ("high_order_surface_flux", C.c_int),
                ("high_order_volume_flux", C.c_int), ("basis", C.c_int), ("entropyproj_limiter", C.c_int),
                ("rhs_limiter", C.c_int), ("bound", C.c_int), ("shockcapture", C.c_int),
                ("upper_bound_mode", C.c_int), ("precision", C.c_int),
                ("shock_a", C.c_double), ("shock_c", C.c_double), ("cell_beta", C.c_double)]


class _Operators(C.Structure):
    _fields_ = [("Nq", C.c_int), ("Nfp", C.c_int), ("Nh", C.c_int), ("Srsh_db", _dp), ("Srs0", _dp),
                ("nnz_h", C.c_int), ("Srsh_nnz", _ip), ("nnz_0", C.c_int), ("Srs0_nnz", _ip), ("Brs", _dp),
                ("MinvVhT", _dp), ("MinvVfT", _dp), ("Vf", _dp), ("Vf_low", _dp), ("wq", _dp), ("fq2q", _ip),
                ("q2fq", _ip), ("VDM_inv", _dp)]


class _Geometry(C.Structure):
    _fields_ = [("Jq", _dp), ("GJh", _dp)]


class _BCData(C.Structure):
    _fields_ = [("mapP", _ip), ("nI", C.c_int), ("mapI", _ip), ("Ival", _dp), ("nO", C.c_int), ("mapO", _ip)]


class PosReport(C.Structure):
    _fields_ = [("violations", C.c_int64), ("first_node", C.c_int64), ("first_stage", C.c_int),
                ("min_rho", C.c_double), ("min_p", C.c_double)]


class RunStats(C.Structure):
    _fields_ = [("steps", C.c_int), ("t", C.c_double), ("last_dt", C.c_double), ("violations", C.c_int64)]


_lib = None


def lib():
    """Load the HIP library; raises if it has not been built (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise P2DEError(P2DE_E_STATE, f"{LIB_PATH} is missing: run __graft_entry__.build()")
        L = C.CDLL(str(LIB_PATH))
        vp = C.c_void_p
        L.p2de_create.argtypes = [C.POINTER(_Config), C.POINTER(_Operators), C.POINTER(_Geometry),
                                  C.POINTER(_BCData), C.POINTER(vp)]
        L.p2de_destroy.argtypes = [vp]
        L.p2de_destroy.restype = None
        L.p2de_set_state.argtypes = [vp, _dp]
        L.p2de_get_state.argtypes = [vp, _dp]
        L.p2de_rhs.argtypes = [vp, C.c_double, C.c_double, C.c_int, _dp, _dp]
        L.p2de_set_debug.argtypes = [vp, C.c_int]
        L.p2de_get_field.argtypes = [vp, C.c_int, _dp]
        L.p2de_ssprk3_step.argtypes = [vp, _dp, _dp]
        L.p2de_run.argtypes = [vp, C.c_int, C.POINTER(RunStats)]
        L.p2de_set_time.argtypes = [vp, C.c_double]
        L.p2de_get_time.argtypes = [vp, _dp, _dp]
        L.p2de_check_positivity.argtypes = [vp, C.POINTER(PosReport)]
        L.p2de_enqueue_steps.argtypes = [vp, C.c_int]
        L.p2de_synchronize.argtypes = [vp]
        L.p2de_set_kernel_timing.argtypes = [vp, C.c_int]
        L.p2de_get_kernel_timing.argtypes = [vp, _dp, _ip]
        L.p2de_num_variants.argtypes = [vp]
        L.p2de_set_variant.argtypes = [vp, C.c_int]
        L.p2de_get_variant.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.p2de_autotune.argtypes = [vp, C.c_int, C.c_char_p, C.POINTER(C.c_int)]
        L.p2de_run_segment.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int64), _dp, C.c_int,
                                       C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.p2de_check_conservation.argtypes = [vp, _dp]
        L.p2de_checkpoint_size.argtypes = [vp]
        L.p2de_checkpoint_size.restype = C.c_int64
        L.p2de_checkpoint_save.argtypes = [vp, C.c_int64, C.c_void_p, C.c_int64]
        L.p2de_checkpoint_load.argtypes = [vp, C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]
        bp = C.c_void_p
        L.p2de_batch_create.argtypes = [C.c_int, C.POINTER(_Config), C.POINTER(_Operators), C.POINTER(_Geometry),
                                        C.POINTER(_BCData), C.POINTER(bp)]
        L.p2de_batch_destroy.argtypes = [bp]
        L.p2de_batch_destroy.restype = None
        L.p2de_batch_num_instances.argtypes = [bp]
        L.p2de_batch_num_dofs.argtypes = [bp, C.c_int]
        L.p2de_batch_num_dofs.restype = C.c_int64
        L.p2de_batch_set_state.argtypes = [bp, C.c_int, _dp]
        L.p2de_batch_get_state.argtypes = [bp, C.c_int, _dp]
        L.p2de_batch_device_state.argtypes = [bp, C.c_int, C.POINTER(vp)]
        L.p2de_batch_set_time.argtypes = [bp, C.c_int, C.c_double]
        L.p2de_batch_get_time.argtypes = [bp, C.c_int, _dp, _dp, C.POINTER(C.c_int)]
        L.p2de_batch_rhs.argtypes = [bp, C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp]
        L.p2de_batch_set_debug.argtypes = [bp, C.c_int]
        L.p2de_batch_get_field.argtypes = [bp, C.c_int, C.c_int, _dp]
        L.p2de_batch_enqueue_steps.argtypes = [bp, C.c_int]
        L.p2de_batch_synchronize.argtypes = [bp]
        L.p2de_batch_run.argtypes = [bp, C.c_int, C.POINTER(C.c_int)]
        L.p2de_batch_check_positivity.argtypes = [bp, C.c_int, C.POINTER(PosReport)]
        L.p2de_batch_set_kernel_timing.argtypes = [bp, C.c_int]
        L.p2de_batch_get_kernel_timing.argtypes = [bp, _dp, _ip]
        L.p2de_batch_set_graph.argtypes = [bp, C.c_int]
        L.p2de_batch_num_variants.argtypes = [bp]
        L.p2de_batch_set_variant.argtypes = [bp, C.c_int]
        L.p2de_batch_get_variant.argtypes = [bp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.p2de_batch_autotune.argtypes = [bp, C.c_int, C.c_char_p, C.POINTER(C.c_int)]
        L.p2de_batch_last_error.argtypes = [bp]
        L.p2de_batch_last_error.restype = C.c_char_p
        L.p2de_num_dofs.argtypes = [vp]
        L.p2de_num_dofs.restype = C.c_int64
        L.p2de_device_state.argtypes = [vp, C.POINTER(vp)]
        L.p2de_last_error.argtypes = [vp]
        L.p2de_last_error.restype = C.c_char_p
        L.p2de_status_string.argtypes = [C.c_int]
        L.p2de_status_string.restype = C.c_char_p
        _lib = L
    return _lib


# every symbol include/p2de_hip.h declares
EXPORTED_SYMBOLS = (
    "p2de_create", "p2de_destroy", "p2de_set_state", "p2de_get_state", "p2de_rhs", "p2de_set_debug",
    "p2de_get_field", "p2de_ssprk3_step", "p2de_run", "p2de_set_time", "p2de_get_time",
    "p2de_check_positivity", "p2de_enqueue_steps", "p2de_synchronize", "p2de_set_kernel_timing",
    "p2de_get_kernel_timing", "p2de_num_dofs", "p2de_device_state", "p2de_last_error", "p2de_status_string",
    "p2de_num_variants", "p2de_set_variant", "p2de_get_variant", "p2de_autotune",
    "p2de_run_segment", "p2de_check_conservation", "p2de_checkpoint_size", "p2de_checkpoint_save",
    "p2de_checkpoint_load",
    "p2de_batch_create", "p2de_batch_destroy", "p2de_batch_num_instances", "p2de_batch_num_dofs",
    "p2de_batch_set_state", "p2de_batch_get_state", "p2de_batch_device_state", "p2de_batch_set_time",
    "p2de_batch_get_time", "p2de_batch_rhs", "p2de_batch_set_debug", "p2de_batch_get_field",
    "p2de_batch_enqueue_steps", "p2de_batch_synchronize", "p2de_batch_run", "p2de_batch_check_positivity",
    "p2de_batch_set_kernel_timing", "p2de_batch_get_kernel_timing", "p2de_batch_set_graph",
    "p2de_batch_num_variants", "p2de_batch_set_variant", "p2de_batch_get_variant", "p2de_batch_autotune",
    "p2de_batch_last_error",
)


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None and a.size else None


def _i(a):
    return a.ctypes.data_as(_ip) if a is not None and a.size else None


def _descriptors(fp: FlatProblem, upper_bound_mode: int = None, precision: int = None):
    """The C-ABI descriptors (p2de_config, p2de_operators, p2de_geometry, p2de_bcdata) of a
    flat problem, plus the arrays they point into (which must outlive the create call)."""
    mode = fp.upper_bound_mode if upper_bound_mode is None else upper_bound_mode
    prec = getattr(fp, "precision", PRECISION_FP64) if precision is None else precision
    cfg = _Config(Nd=fp.Nd, N=fp.N1D - 1, Kx=fp.Kx, Ky=fp.Ky, gamma=fp.gamma, zeta=fp.zeta, eta=fp.eta,
                  POSTOL=fp.POSTOL, ZEROTOL=fp.ZEROTOL, CFL=fp.CFL, dt0=fp.dt0, t0=fp.t0, T=fp.T,
                  high_order_surface_flux=1, upper_bound_mode=int(mode), precision=int(prec),
                  low_order_surface_flux=int(getattr(fp, "low_order_surface_flux", 0)),
                  basis=int(getattr(fp, "basis", 0)), entropyproj_limiter=int(getattr(fp, "entropyproj_limiter", 0)),
                  rhs_limiter=int(getattr(fp, "rhs_limiter", 0)), bound=int(getattr(fp, "bound", 0)),
                  shockcapture=int(getattr(fp, "shockcapture", 0)), shock_a=float(getattr(fp, "shock_a", 0.5)),
                  shock_c=float(getattr(fp, "shock_c", 1.8)), cell_beta=float(getattr(fp, "cell_beta", 0.5)))
    keep = {}
    for n in ("Srsh_db", "Srs0", "Brs", "MinvVhT", "MinvVfT", "Vf", "Vf_low", "wq", "Jq", "GJh", "Ival"):
        keep[n] = np.ascontiguousarray(getattr(fp, n), dtype=np.float64)
    vinv = getattr(fp, "VDM_inv", None)
    keep["VDM_inv"] = None if vinv is None else np.ascontiguousarray(vinv, dtype=np.float64)
    for n in ("Srsh_nnz", "Srs0_nnz", "fq2q", "q2fq", "mapP", "mapI", "mapO"):
        keep[n] = np.ascontiguousarray(getattr(fp, n), dtype=np.int64)
    ops = _Operators(Nq=fp.Nq, Nfp=fp.Nfp, Nh=fp.Nh, Srsh_db=_d(keep["Srsh_db"]), Srs0=_d(keep["Srs0"]),
                     nnz_h=len(keep["Srsh_nnz"]) // 2, Srsh_nnz=_i(keep["Srsh_nnz"]),
                     nnz_0=len(keep["Srs0_nnz"]) // 2, Srs0_nnz=_i(keep["Srs0_nnz"]), Brs=_d(keep["Brs"]),
                     MinvVhT=_d(keep["MinvVhT"]), MinvVfT=_d(keep["MinvVfT"]), Vf=_d(keep["Vf"]),
                     Vf_low=_d(keep["Vf_low"]), wq=_d(keep["wq"]), fq2q=_i(keep["fq2q"]), q2fq=_i(keep["q2fq"]),
                     VDM_inv=_d(keep["VDM_inv"]) if keep["VDM_inv"] is not None else None)
    geo = _Geometry(Jq=_d(keep["Jq"]), GJh=_d(keep["GJh"]))
    bc = _BCData(mapP=_i(keep["mapP"]), nI=len(keep["mapI"]), mapI=_i(keep["mapI"]), Ival=_d(keep["Ival"]),
                 nO=len(keep["mapO"]), mapO=_i(keep["mapO"]))
    return cfg, ops, geo, bc, keep


class DeviceSolver:
    """One native context; the state lives on the GPU between calls."""

    def __init__(self, fp: FlatProblem, upper_bound_mode: int = None, precision: int = None):
        self.fp = fp
        self.Nc, self.Nq, self.K = fp.Nc, fp.Nq, fp.K
        cfg, ops, geo, bc, self._keep = _descriptors(fp, upper_bound_mode, precision)
        h = C.c_void_p()
        st = lib().p2de_create(C.byref(cfg), C.byref(ops), C.byref(geo), C.byref(bc), C.byref(h))
        if st != P2DE_OK:
            raise P2DEError(st, lib().p2de_last_error(None).decode())
        self.h = h

    # --------------------------------------------------------------- plumbing
    def _check(self, st):
        if st != P2DE_OK:
            raise P2DEError(st, lib().p2de_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().p2de_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        return (self.K, self.Nq, self.Nc)

    # --------------------------------------------------------------- state
    def set_state(self, Uq: np.ndarray):
        U = np.ascontiguousarray(Uq, dtype=np.float64)
        if U.size != self.K * self.Nq * self.Nc:
            raise P2DEError(P2DE_E_INVALID, f"state has {U.size} values, expected {self.K * self.Nq * self.Nc}")
        self._check(lib().p2de_set_state(self.h, _d(U)))

    def get_state(self) -> np.ndarray:
        out = np.empty(self.shape)
        self._check(lib().p2de_get_state(self.h, _d(out)))
        return out

    # --------------------------------------------------------------- rhs!
    def rhs(self, t: float, dt: float, nstage: int, want_rhs: bool = True):
        """``rhs!`` on the device state; returns (dt_out, rhsU or None)."""
        out = np.empty(self.shape) if want_rhs else None
        dto = C.c_double(0.0)
        self._check(lib().p2de_rhs(self.h, float(t), float(dt), int(nstage), C.byref(dto), _d(out) if want_rhs else None))
        return dto.value, out

    def set_debug(self, on: bool = True):
        self._check(lib().p2de_set_debug(self.h, 1 if on else 0))

    def field(self, which: int) -> np.ndarray:
        n = self.K * self.Nq * self.Nc
        if which == FIELD_L_LOCAL:
            n = (self.Nq + self.fp.N1D) * self.fp.Nd * self.K
        elif which == FIELD_THETA:
            n = self.fp.Nfp * self.K
        elif which in (FIELD_L, FIELD_THETA_MEAN):
            n = 3 * self.K
        out = np.empty(n)
        self._check(lib().p2de_get_field(self.h, int(which), _d(out)))
        if which in (FIELD_L, FIELD_THETA_MEAN):
            return out.reshape(3, self.K).T      # (K, 3) as the reference's L and theta
        return out

    # --------------------------------------------------------------- time stepping
    def step(self, t: float):
        """One SSP33! step of the device state; returns (t_new, dt)."""
        tt = C.c_double(t)
        dt = C.c_double(0.0)
        self._check(lib().p2de_ssprk3_step(self.h, C.byref(tt), C.byref(dt)))
        return tt.value, dt.value

    def set_time(self, t: float):
        self._check(lib().p2de_set_time(self.h, float(t)))

    def get_time(self):
        t, dt = C.c_double(0.0), C.c_double(0.0)
        self._check(lib().p2de_get_time(self.h, C.byref(t), C.byref(dt)))
        return t.value, dt.value

    def run(self, max_steps: int = 0) -> RunStats:
        st = RunStats()
        self._check(lib().p2de_run(self.h, int(max_steps), C.byref(st)))
        return st

    def run_segment(self, output_interval: int, step_index: int, max_steps: int = 0, cap: int = 4096):
        """Steps until the next history entry of SSP33!'s loop (``p2de_run_segment``); returns
        (dts of the steps taken, new step index, whether the segment ended at a history entry)."""
        dts = np.empty(max(1, int(cap)))
        i = C.c_int64(int(step_index))
        n, out = C.c_int(0), C.c_int(0)
        self._check(lib().p2de_run_segment(self.h, int(output_interval), int(max_steps), C.byref(i), _d(dts),
                                           int(dts.size), C.byref(n), C.byref(out)))
        return dts[:n.value].copy(), int(i.value), bool(out.value)

    def check_conservation(self) -> float:
        """``check_conservation`` (src/dg/utils.jl:1-12) of the device state."""
        tot = C.c_double(0.0)
        self._check(lib().p2de_check_conservation(self.h, C.byref(tot)))
        return tot.value

    def checkpoint(self, step_index: int = 1) -> bytes:
        """The context's resumable snapshot (``p2de_checkpoint_save``) as bytes."""
        n = int(lib().p2de_checkpoint_size(self.h))
        buf = C.create_string_buffer(n)
        self._check(lib().p2de_checkpoint_save(self.h, int(step_index), buf, n))
        return buf.raw

    def restore(self, blob: bytes) -> int:
        """Loads a snapshot taken by ``checkpoint``; returns its step index."""
        i = C.c_int64(0)
        buf = C.create_string_buffer(bytes(blob), len(blob))
        self._check(lib().p2de_checkpoint_load(self.h, buf, len(blob), C.byref(i)))
        return int(i.value)

    def enqueue_steps(self, n: int):
        self._check(lib().p2de_enqueue_steps(self.h, int(n)))

    def synchronize(self):
        self._check(lib().p2de_synchronize(self.h))

    def check_positivity(self) -> PosReport:
        r = PosReport()
        self._check(lib().p2de_check_positivity(self.h, C.byref(r)))
        return r

    def set_kernel_timing(self, on: bool):
        self._check(lib().p2de_set_kernel_timing(self.h, 1 if on else 0))

    def kernel_timing(self):
        ms = (C.c_double * 3)()
        cnt = (C.c_int64 * 3)()
        self._check(lib().p2de_get_kernel_timing(self.h, ms, cnt))
        return list(ms), list(cnt)

    def num_variants(self) -> int:
        return int(lib().p2de_num_variants(self.h))

    def set_variant(self, v: int):
        self._check(lib().p2de_set_variant(self.h, int(v)))

    def variant(self):
        v, e, nt = C.c_int(), C.c_int(), C.c_int()
        self._check(lib().p2de_get_variant(self.h, C.byref(v), C.byref(e), C.byref(nt)))
        return v.value, e.value, nt.value

    def autotune(self, steps: int = 5, cache_path: str = None) -> int:
        ch = C.c_int(-1)
        self._check(lib().p2de_autotune(self.h, int(steps), cache_path.encode() if cache_path else None, C.byref(ch)))
        return ch.value

    def num_dofs(self) -> int:
        return int(lib().p2de_num_dofs(self.h))


class BatchSolver:
    """Many independent 2D solves in mixed precision, advanced together (``p2de_batch_*``).
    Instance ``i`` is ``fps[i]``; its state has the shape ``(K_i, Nq_i, 4)``."""

    def __init__(self, fps, upper_bound_mode: int = None):
        n = len(fps)
        if n < 1:
            raise P2DEError(P2DE_E_INVALID, "empty batch")
        self.fps = list(fps)
        cfgs, opss, geos, bcs = (_Config * n)(), (_Operators * n)(), (_Geometry * n)(), (_BCData * n)()
        self._keep = []
        for i, fp in enumerate(self.fps):
            cfg, ops, geo, bc, keep = _descriptors(fp, upper_bound_mode, PRECISION_MIXED)
            cfgs[i], opss[i], geos[i], bcs[i] = cfg, ops, geo, bc
            self._keep.append(keep)
        h = C.c_void_p()
        st = lib().p2de_batch_create(n, cfgs, opss, geos, bcs, C.byref(h))
        if st != P2DE_OK:
            raise P2DEError(st, lib().p2de_batch_last_error(None).decode())
        self.h = h
        self.n = n

    def _check(self, st):
        if st != P2DE_OK:
            raise P2DEError(st, lib().p2de_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().p2de_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, i: int):
        fp = self.fps[i]
        return (fp.K, fp.Nq, fp.Nc)

    def num_dofs(self, i: int = -1) -> int:
        return int(lib().p2de_batch_num_dofs(self.h, int(i)))

    def set_state(self, i: int, Uq: np.ndarray):
        U = np.ascontiguousarray(Uq, dtype=np.float64)
        if U.size != int(np.prod(self.shape(i))):
            raise P2DEError(P2DE_E_INVALID, f"state of instance {i} has {U.size} values")
        self._check(lib().p2de_batch_set_state(self.h, int(i), _d(U)))

    def get_state(self, i: int) -> np.ndarray:
        out = np.empty(self.shape(i))
        self._check(lib().p2de_batch_get_state(self.h, int(i), _d(out)))
        return out

    def set_time(self, t: float, i: int = -1):
        self._check(lib().p2de_batch_set_time(self.h, int(i), float(t)))

    def get_time(self, i: int):
        t, dt, n = C.c_double(0.0), C.c_double(0.0), C.c_int(0)
        self._check(lib().p2de_batch_get_time(self.h, int(i), C.byref(t), C.byref(dt), C.byref(n)))
        return t.value, dt.value, n.value

    def rhs(self, i: int, t: float, dt: float, nstage: int, want_rhs: bool = True):
        out = np.empty(self.shape(i)) if want_rhs else None
        dto = C.c_double(0.0)
        self._check(lib().p2de_batch_rhs(self.h, int(i), float(t), float(dt), int(nstage), C.byref(dto),
                                         _d(out) if want_rhs else None))
        return dto.value, out

    def set_debug(self, on: bool = True):
        self._check(lib().p2de_batch_set_debug(self.h, 1 if on else 0))

    def field(self, i: int, which: int) -> np.ndarray:
        out = np.empty(self.shape(i))
        self._check(lib().p2de_batch_get_field(self.h, int(i), int(which), _d(out)))
        return out

    def enqueue_steps(self, n: int):
        self._check(lib().p2de_batch_enqueue_steps(self.h, int(n)))

    def synchronize(self):
        self._check(lib().p2de_batch_synchronize(self.h))

    def run(self, max_steps: int = 0) -> int:
        n = C.c_int(0)
        self._check(lib().p2de_batch_run(self.h, int(max_steps), C.byref(n)))
        return n.value

    def check_positivity(self, i: int) -> PosReport:
        r = PosReport()
        self._check(lib().p2de_batch_check_positivity(self.h, int(i), C.byref(r)))
        return r

    def set_kernel_timing(self, on: bool):
        self._check(lib().p2de_batch_set_kernel_timing(self.h, 1 if on else 0))

    def kernel_timing(self):
        ms = (C.c_double * 3)()
        cnt = (C.c_int64 * 3)()
        self._check(lib().p2de_batch_get_kernel_timing(self.h, ms, cnt))
        return list(ms), list(cnt)

    def set_graph(self, on: bool):
        self._check(lib().p2de_batch_set_graph(self.h, 1 if on else 0))

    def num_variants(self) -> int:
        return int(lib().p2de_batch_num_variants(self.h))

    def set_variant(self, v: int):
        self._check(lib().p2de_batch_set_variant(self.h, int(v)))

    def variant(self):
        v, nt = C.c_int(), C.c_int()
        E = (C.c_int * 6)()
        self._check(lib().p2de_batch_get_variant(self.h, C.byref(v), C.byref(nt), E))
        return v.value, nt.value, list(E)

    def autotune(self, steps: int = 3, cache_path: str = None) -> int:
        ch = C.c_int(-1)
        self._check(lib().p2de_batch_autotune(self.h, int(steps), cache_path.encode() if cache_path else None,
                                              C.byref(ch)))
        return ch.value


# ------------------------------------------------------------------- reference-shaped API
def _device(state: State, solver: Solver, state_param: StateParam) -> DeviceSolver:
    if state.device is None:
        state.device = DeviceSolver(flat_problem(solver, state_param))
    return state.device


def rhs(state: State, solver: Solver, state_param: StateParam, time_param: TimeParam) -> float:
    """``rhs!`` (src/dg/rhs/rhs.jl:5-13): evaluates the limited RHS of
    ``state.preallocation.Uq`` on the GPU, writes ``state.preallocation.rhsU`` and returns
    the CFL dt at stage 1 (``time_param.dt`` otherwise)."""
    dev = _device(state, solver, state_param)
    dev.set_state(state.preallocation.Uq)
    dt, r = dev.rhs(time_param.t, time_param.dt, time_param.nstage)
    state.preallocation.rhsU[...] = r
    return dt


@dataclass
class DataHistory:
    """``DataHistory`` (src/dg/PostprocessData.jl:1-7): the states, element limiting
    parameters ``L``, element entropy-projection parameters ``theta`` and times of the history
    entries, every step's dt, plus the conservation totals printed at each entry and the
    positivity violations the device counted."""
    Uhist: list = field(default_factory=list)
    Lhist: list = field(default_factory=list)
    thetahist: list = field(default_factory=list)
    thist: list = field(default_factory=list)
    dthist: list = field(default_factory=list)
    conservation: list = field(default_factory=list)
    positivity_violations: int = 0


def SSP33(state: State, solver: Solver, state_param: StateParam, max_steps: int = 0, verbose: bool = False,
          resume: bytes = None) -> DataHistory:
    """``SSP33!`` (src/timestepping/SSPRK33.jl:1-62): advances ``state.preallocation.Uq``
    from t0 to T on the GPU and returns its ``DataHistory``.  Between history entries
    (``mod(i, output_interval) == 0`` or ``|t - T| < 1e-10``, SSPRK33.jl:46) the steps stay
    queued on the device; at each entry the state, ``L``, ``theta`` and time are copied out and
    ``check_conservation`` taken, as the reference does.  ``resume``: a ``DeviceSolver.checkpoint``
    blob to continue from instead of (Uq, t0); the history then starts at the checkpoint.
    ``max_steps > 0`` stops early (the state is then the one reached)."""
    T = solver.param.timestepping_param.T
    interval = int(solver.param.postprocessing_param.output_interval)
    dev = _device(state, solver, state_param)
    hist = DataHistory()
    if resume is not None:
        i = dev.restore(resume)
    else:
        dev.set_state(state.preallocation.Uq)
        dev.set_time(solver.param.timestepping_param.t0)
        i = 1
    taken = 0
    while max_steps <= 0 or taken < max_steps:
        left = max_steps - taken if max_steps > 0 else 0
        dts, i, at_output = dev.run_segment(interval, i, max_steps=left)
        hist.dthist.extend(float(d) for d in dts)
        taken += len(dts)
        if at_output:
            t, dt = dev.get_time()
            hist.thist.append(t)
            hist.Uhist.append(dev.get_state())
            hist.Lhist.append(dev.field(FIELD_L))
            hist.thetahist.append(dev.field(FIELD_THETA_MEAN))
            total = dev.check_conservation()
            hist.conservation.append(total)
            if verbose:
                print(f"Current time {t} with time step size {dt}, and final time {T}, step {i}", flush=True)
                print(f"total_conservation = {total}", flush=True)
        if len(dts) == 0 or not dev.get_time()[0] < T:
            break
    state.preallocation.Uq[...] = dev.get_state()
    hist.positivity_violations = int(dev.check_positivity().violations)
    return hist
