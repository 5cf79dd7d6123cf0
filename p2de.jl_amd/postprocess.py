"""Host-side output around the GPU path: error norms, the error table, the history file,
VTK output and checkpoint files.

* ``calculate_error`` follows src/dg/postprocess.jl:1-46;
* ``check_conservation`` (host arrays) follows src/dg/utils.jl:1-12; the device form is
  ``DeviceSolver.check_conservation``;
* ``write_error_table``/``visualize_error_data`` keep the rows that write_to_jld2 and
  visualize_error_data keep (postprocess.jl:131-146) in a CSV file instead of a JLD2
  DataFrame, and ``save_history`` the ``DataHistory`` arrays in an ``.npz``;
* ``postprocessing_cache``/``construct_vtk_file`` follow postprocess.jl:148-213 and write VTK
  XML structured grids (``.vts``) plus a ParaView collection (``.pvd``) without WriteVTK;
* ``save_checkpoint``/``load_checkpoint`` store the library's checkpoint block in a file.
"""
from __future__ import annotations

import base64
import csv
import os
from dataclasses import dataclass
from pathlib import Path
from typing import Callable, Optional

import numpy as np

from .dg import Solver, primitive_to_conservative


@dataclass(frozen=True)
class ErrorData:
    """``ErrorData`` (src/common/types/PostprocessData.jl:9-13)."""
    L1err: float
    L2err: float
    Linferr: float


def exact_state(solver: Solver, exact_sol: Callable, T: float) -> np.ndarray:
    """``exact_solution`` (postprocess.jl:48-59) at every node: ``exact_sol(equation, x[, y], t)``
    returns primitive variables, as the reference's examples define it; result (K, Nq, Nc)."""
    md = solver.md
    eq = solver.param.equation
    if eq.Nd == 1:
        prim = exact_sol(eq, md.xq, T)
    else:
        prim = exact_sol(eq, md.xq, md.yq, T)
    prim = [np.broadcast_to(np.asarray(p, dtype=np.float64), md.xq.shape) for p in prim]
    U = primitive_to_conservative(eq.gamma, prim)              # (Nc, Nq, K)
    return np.ascontiguousarray(np.transpose(U, (2, 1, 0)))


def calculate_error(Uq: np.ndarray, solver: Solver, exact_sol: Callable, T: Optional[float] = None,
                    verbose: bool = False) -> ErrorData:
    """``calculate_error`` (postprocess.jl:1-46): per component the wq*J-weighted L1 and L2
    norms and the max norm of U - U_exact at T, each divided by the same norm of U_exact; the
    three results sum these ratios over the components whose exact max norm exceeds 1e-14.

    The weight of node i is ``wq[i] * Jq[i]`` with the linear index i (postprocess.jl:18), i.e.
    the first element's Jacobian for every element; on the uniform meshes the reference
    builds all elements have that Jacobian."""
    if T is None:
        T = solver.param.timestepping_param.T
    U = np.asarray(Uq, dtype=np.float64)
    K, Nq, Nc = U.shape
    Ue = exact_state(solver, exact_sol, T)
    wJ = (solver.discrete_data.ops.wq * solver.md.J[:, 0])[None, :, None]     # (1, Nq, 1)
    d = np.abs(Ue - U)
    a = np.abs(Ue)
    L1err = np.sum(wJ * d, axis=(0, 1))
    L2err = np.sum(wJ * d ** 2, axis=(0, 1))
    Linferr = np.max(d, axis=(0, 1))
    L1exact = np.sum(wJ * a, axis=(0, 1))
    L2exact = np.sum(wJ * a ** 2, axis=(0, 1))
    Linfexact = np.max(a, axis=(0, 1))
    L1 = L2 = Linf = 0.0
    for c in range(Nc):
        if Linfexact[c] > 1e-14:
            L1 += L1err[c] / L1exact[c]
            L2 += np.sqrt(L2err[c]) / np.sqrt(L2exact[c])
            Linf += Linferr[c] / Linfexact[c]
    err = ErrorData(float(L1), float(L2), float(Linf))
    if verbose:
        print(f"N = {solver.param.N}, K = {solver.param.K}")
        print(f"L1 error is {err.L1err}")
        print(f"L2 error is {err.L2err}")
        print(f"Linf error is {err.Linferr}")
    return err


def check_conservation(Uq: np.ndarray, solver: Solver) -> float:
    """``check_conservation`` (src/dg/utils.jl:1-12) of a host state: the sum over the components
    of sum_k sum_i J[i,k] wq[i] Uq[i,k]."""
    U = np.asarray(Uq, dtype=np.float64)
    w = solver.md.J.T * solver.discrete_data.ops.wq[None, :]            # (K, Nq)
    return float(np.sum(np.sum(w[:, :, None] * U, axis=(0, 1))))


# ------------------------------------------------------------------------- error table
TABLE_COLUMNS = ("N", "K", "timestepping_param", "limiting_param", "approximation_basis", "rhs",
                 "entropyproj_limiter", "rhs_limiter", "L1err", "L2err", "Linferr", "num_steps")


def error_table_row(param, data_hist, err: ErrorData) -> dict:
    """The columns visualize_error_data shows (postprocess.jl:144-146) for one run."""
    return {"N": param.N, "K": param.K, "timestepping_param": repr(param.timestepping_param),
            "limiting_param": repr(param.limiting_param),
            "approximation_basis": type(param.approximation_basis).__name__, "rhs": repr(param.rhs),
            "entropyproj_limiter": type(param.entropyproj_limiter).__name__, "rhs_limiter": repr(param.rhs_limiter),
            "L1err": repr(err.L1err), "L2err": repr(err.L2err), "Linferr": repr(err.Linferr),
            "num_steps": len(data_hist.dthist)}


def write_error_table(param, data_hist, err: ErrorData, path, show: bool = True):
    """write_to_jld2 (postprocess.jl:131-142): appends the run's row to the table at ``path``
    (CSV, created with a header when absent), then prints the table."""
    path = Path(path)
    new = not path.exists()
    with open(path, "a", newline="") as f:
        w = csv.DictWriter(f, fieldnames=TABLE_COLUMNS)
        if new:
            w.writeheader()
        w.writerow(error_table_row(param, data_hist, err))
    if show:
        visualize_error_data(path)


def read_error_table(path) -> list:
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def visualize_error_data(path) -> str:
    """visualize_error_data (postprocess.jl:144-146): the table as aligned text (printed)."""
    rows = read_error_table(path)
    cols = list(TABLE_COLUMNS)
    width = {c: max([len(c)] + [len(r[c]) for r in rows]) for c in cols}
    lines = [" | ".join(c.ljust(width[c]) for c in cols), "-+-".join("-" * width[c] for c in cols)]
    lines += [" | ".join(r[c].ljust(width[c]) for c in cols) for r in rows]
    text = "\n".join(lines)
    print(text)
    return text


def save_history(path, data_hist):
    """The ``DataHistory`` arrays (PostprocessData.jl:1-7) in an ``.npz`` (numbers only)."""
    U = np.stack(data_hist.Uhist) if data_hist.Uhist else np.zeros((0,))
    L = np.stack(data_hist.Lhist) if data_hist.Lhist else np.zeros((0,))
    th = np.stack(data_hist.thetahist) if data_hist.thetahist else np.zeros((0,))
    np.savez(path, Uhist=U, Lhist=L, thetahist=th, thist=np.asarray(data_hist.thist, dtype=np.float64),
             dthist=np.asarray(data_hist.dthist, dtype=np.float64),
             conservation=np.asarray(getattr(data_hist, "conservation", []), dtype=np.float64))


def load_history(path) -> dict:
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


# ------------------------------------------------------------------------- VTK output
@dataclass
class PostprocessingCache:
    """``PostprocessingCache``: the plot grid of a rectangular 2D mesh (postprocess.jl:160-182)."""
    xp: np.ndarray
    yp: np.ndarray
    Up: np.ndarray


def _element_blocks(solver: Solver):
    Kx, Ky = solver.param.K
    N1D = solver.param.N + 1
    for k in range(Kx * Ky):
        ik, jk = k % Kx, k // Kx                        # mod1(k, Kx), div(k - 1, Kx) + 1
        yield k, slice(ik * N1D, (ik + 1) * N1D), slice(jk * N1D, (jk + 1) * N1D)


def postprocessing_cache(solver: Solver) -> PostprocessingCache:
    """postprocess.jl:160-182: node coordinates on an (N1D*Kx, N1D*Ky) grid, element k's nodes
    reshaped column-major into its N1D x N1D block."""
    if solver.param.equation.Nd != 2:
        raise ValueError("VTK output is written for rectangular 2D quad meshes")
    Kx, Ky = solver.param.K
    N1D = solver.param.N + 1
    Nc = solver.param.equation.Nc
    xp = np.zeros((N1D * Kx, N1D * Ky))
    yp = np.zeros_like(xp)
    for k, ir, jr in _element_blocks(solver):
        xp[ir, jr] = solver.md.xq[:, k].reshape(N1D, N1D, order="F")
        yp[ir, jr] = solver.md.yq[:, k].reshape(N1D, N1D, order="F")
    return PostprocessingCache(xp=xp, yp=yp, Up=np.zeros(xp.shape + (Nc,)))


def fill_plot_grid(cache: PostprocessingCache, solver: Solver, U: np.ndarray) -> np.ndarray:
    N1D = solver.param.N + 1
    for k, ir, jr in _element_blocks(solver):
        for c in range(U.shape[2]):
            cache.Up[ir, jr, c] = U[k, :, c].reshape(N1D, N1D, order="F")
    return cache.Up


def _b64(a: np.ndarray) -> str:
    """VTK XML inline binary: base64 of a UInt64 byte count followed by the little-endian data."""
    raw = np.ascontiguousarray(a).astype("<f8").tobytes()
    return base64.b64encode(np.uint64(len(raw)).astype("<u8").tobytes() + raw).decode("ascii")


def write_vts(path, xp: np.ndarray, yp: np.ndarray, fields: dict):
    """A VTK XML StructuredGrid with points (x, y, 0) and Float64 point fields; grid index i
    (first axis) runs fastest, as WriteVTK orders a Julia matrix."""
    ni, nj = xp.shape
    pts = np.stack([xp.T.ravel(), yp.T.ravel(), np.zeros(ni * nj)], axis=1)
    ext = f"0 {ni - 1} 0 {nj - 1} 0 0"
    out = ['<?xml version="1.0"?>',
           '<VTKFile type="StructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">',
           f'  <StructuredGrid WholeExtent="{ext}">', f'    <Piece Extent="{ext}">',
           '      <Points>',
           f'        <DataArray type="Float64" NumberOfComponents="3" format="binary">{_b64(pts)}</DataArray>',
           '      </Points>', '      <PointData>']
    for name, a in fields.items():
        out.append(f'        <DataArray type="Float64" Name="{name}" format="binary">{_b64(np.asarray(a).T.ravel())}'
                   '</DataArray>')
    out += ['      </PointData>', '    </Piece>', '  </StructuredGrid>', '</VTKFile>']
    Path(path).write_text("\n".join(out) + "\n")


def construct_vtk_file(cache: PostprocessingCache, solver: Solver, data_hist, output_path, filename) -> Path:
    """construct_vtk_file! (postprocess.jl:184-207): one ``.vts`` per history entry with rho,
    rhou, rhov, E (set_vtk_field!, postprocess.jl:209-214) and a ``.pvd`` collection keyed by
    time.  Returns the ``.pvd`` path."""
    out = Path(output_path)
    out.mkdir(parents=True, exist_ok=True)
    N = solver.param.N
    Kx, Ky = solver.param.K
    stem = f"{filename}_N={N}_K=({Kx}, {Ky})"
    entries = []
    for U, t in zip(data_hist.Uhist, data_hist.thist):
        Up = fill_plot_grid(cache, solver, U)
        name = f"{stem}_t={t}.vts"
        write_vts(out / name, cache.xp, cache.yp,
                  {"rho": Up[..., 0], "rhou": Up[..., 1], "rhov": Up[..., 2], "E": Up[..., 3]})
        entries.append((t, name))
    pvd = out / f"{stem}.pvd"
    lines = ['<?xml version="1.0"?>', '<VTKFile type="Collection" version="1.0" byte_order="LittleEndian">',
             '  <Collection>']
    lines += [f'    <DataSet timestep="{t!r}" part="0" file="{name}"/>' for t, name in entries]
    lines += ['  </Collection>', '</VTKFile>']
    pvd.write_text("\n".join(lines) + "\n")
    return pvd


def read_vts(path) -> dict:
    """Points and point fields of a file ``write_vts`` wrote, as (ni, nj)-shaped arrays."""
    import xml.etree.ElementTree as ET
    root = ET.parse(path).getroot()
    piece = root.find("StructuredGrid/Piece")
    e = [int(v) for v in piece.get("Extent").split()]
    ni, nj = e[1] - e[0] + 1, e[3] - e[2] + 1

    def dec(node):
        raw = base64.b64decode(node.text)
        n = int(np.frombuffer(raw[:8], dtype="<u8")[0])
        return np.frombuffer(raw[8:8 + n], dtype="<f8")

    pts = dec(piece.find("Points/DataArray")).reshape(nj, ni, 3)
    out = {"x": pts[..., 0].T, "y": pts[..., 1].T}
    for da in piece.find("PointData"):
        out[da.get("Name")] = dec(da).reshape(nj, ni).T
    return out


# ------------------------------------------------------------------------- checkpoints
def save_checkpoint(dev, path, step_index: int = 1):
    """Writes ``dev``'s checkpoint block (``p2de_checkpoint_save``) to ``path``; the file is
    written beside it first and renamed, so a crash leaves the previous checkpoint whole."""
    blob = dev.checkpoint(step_index)
    path = Path(path)
    tmp = path.with_name(path.name + ".tmp")
    with open(tmp, "wb") as f:
        f.write(blob)
        f.flush()
        os.fsync(f.fileno())
    os.replace(tmp, path)


def load_checkpoint(dev, path) -> int:
    """Loads a checkpoint file into ``dev``; returns its step index."""
    return dev.restore(Path(path).read_bytes())
