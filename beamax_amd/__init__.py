"""beamax_amd: MI355X-native MSGB forward/adjoint operator (HIP, gfx950).

Drop-in for the MSGB path of elma16/beamax: ``Domain``, ``Sensor``,
``DyadicDecomposition``, ``MSWPT`` and ``solvers.MSGBSolver`` keep the
reference API; the numerics run as hand-written HIP kernels behind the C-ABI
in ``include/beamax_amd.h`` (``libbeamax_amd.so``, loaded with ctypes).
"""

from .decomposition import DyadicDecomposition
from .geometry import Domain, Sensor
from .transforms import MSWPT
from . import gb, solvers, utils
from .solvers import MSGBSolver

__version__ = "0.1.0"

__all__ = ["Domain", "Sensor", "DyadicDecomposition", "MSWPT", "MSGBSolver", "gb", "solvers",
           "utils"]
