"""miniaero_amd -- MI355X-native explicit RK4 residual path of miniAero.

The numerical work lives in hand-written HIP kernels (libminiaero_amd.so,
C ABI declared in include/miniaero.h).  This package is the Python host
mirror of the reference's interfaces:

    Options                 Options.h (miniaero.inp)
    Parallel3DMesh          Parallel3DMesh.h  -> MeshData
    TimeSolverExplicitRK4   TimeSolverExplicitRK4.h (Solve)
    results / numeric_text_diff  results.<rank> files and their comparator
"""
from ._lib import MiniAeroError, load
from .mesh import Faces, MeshData, Parallel3DMesh, plan_batch
from .options import Options
from .results import numeric_file_diff, numeric_text_diff, read_results, write_results
from .server import Server
from .solver import BatchSolver, TimeSolverExplicitRK4

__all__ = [
    "MiniAeroError", "load", "Faces", "MeshData", "Parallel3DMesh", "Options",
    "TimeSolverExplicitRK4", "BatchSolver", "Server", "numeric_text_diff", "numeric_file_diff", "read_results",
    "write_results",
]
