"""Writes tests/golden/weno5_shuosher.npz from the reference's data/weno5_shuosher.mat: the
WENO5 density of the Shu-Osher problem at T = 1.8 on 25002 points of [-5, 5], which
examples/1D/shu-osher.jl:70-74 plots its solution against.  Every 5th point is kept (x and
rho), enough to interpolate the solution at the nodes of a 64..256-element mesh.

Run in a container that has the reference checkout:
    python tests/golden/make_weno5_shuosher.py [/path/to/reference]
The .mat file is MATLAB v5 numeric data; scipy.io.loadmat reads it without executing anything.
"""
import sys
from pathlib import Path

import numpy as np
import scipy.io

ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
m = scipy.io.loadmat(str(ref / "data" / "weno5_shuosher.mat"))
x = np.asarray(m["x"], dtype=np.float64).ravel()
rho = np.asarray(m["rho"], dtype=np.float64).ravel()
keep = np.r_[np.arange(0, x.size, 5), x.size - 1]
keep = np.unique(keep)
out = Path(__file__).resolve().parent / "weno5_shuosher.npz"
np.savez_compressed(out, x=x[keep], rho=rho[keep], T=np.float64(1.8), source_points=np.int64(x.size))
print(f"wrote {out}: {keep.size} of {x.size} points")
