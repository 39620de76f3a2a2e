"""diffnano_amd -- MI355X-native differentiable Yee FDTD.

Drop-in for DiffNano's ``FDTDSolver2D`` / ``FDTDSolver3D`` (and its
``SimResult`` / ``Solver`` protocol); the time stepping, DFT monitors and
the checkpoint-and-replay adjoint run in hand-written gfx950 HIP kernels
(``libdnx_fdtd.so``, C ABI in ``include/dnx_fdtd.h``).
"""

from ._result import SimResult, Solver
from .fdtd2d import FDTDSolver2D
from .fdtd3d import FDTDSolver3D

__all__ = ["FDTDSolver2D", "FDTDSolver3D", "SimResult", "Solver"]
__version__ = "0.1.0"
