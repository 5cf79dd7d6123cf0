"""p2de.jl_amd -- MI355X-native drop-in for P2DE.jl's limited ESDG right-hand side and
SSP-RK3 stage update (1D/2D compressible Euler).

Host-side mirror of the reference interface:

* option and parameter types (``types``), named as in src/common/types/Solver.jl;
* ``initialize_DG`` and the state containers (``dg``), as in src/dg/init.jl:44-60;
* ``rhs`` / ``SSP33`` (``solver``), the drop-ins for ``rhs!`` (src/dg/rhs/rhs.jl:5-13)
  and ``SSP33!`` (src/timestepping/SSPRK33.jl:1-62), which run on the GPU through the
  C-ABI library ``libp2de_hip.so`` (``include/p2de_hip.h``);
* ``BatchSolver`` (``solver``), many independent 2D solves advanced together in mixed
  precision (FP32 bulk, FP64 limiter and RK stages), and the batched workload
  (``configs.batched_instances``);
* ``postprocess``: ``calculate_error`` (src/dg/postprocess.jl:1-46), ``check_conservation``,
  the error table and history files, VTK output and checkpoint files.

Import it with ``__graft_entry__.load_package()``: the directory name is not a Python
identifier, so it is registered as ``p2de_jl_amd``.
"""
from . import types, operators, dg, configs, timestep  # noqa: F401
from .types import *  # noqa: F401,F403
from .dg import (initialize_DG, make_periodic, primitive_to_conservative, flat_problem,  # noqa: F401
                 Solver, State, StateParam, TimeParam)
from . import solver  # noqa: F401,E402
from .solver import BatchSolver, DataHistory, DeviceSolver, P2DEError, SSP33, rhs  # noqa: F401,E402
from . import postprocess  # noqa: F401,E402
from .postprocess import ErrorData, calculate_error, check_conservation  # noqa: F401,E402
