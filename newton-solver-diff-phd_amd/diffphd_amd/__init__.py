"""MI355X-native DiffPhD projective-dynamics solver path.

Public names mirror the reference package (``src/newton_solver_diff_phd/__init__.py:1-11``):
``SolverDiffPhD`` and ``Mesh``, plus the Newton-shaped model holders this
package needs without Newton.
"""

from ._lib import NativeLibraryMissing
from .model import Control, Mesh, Model, ModelBuilder, State
from .solver import ConvergenceError, SolverDiffPhD

__all__ = [
    "ConvergenceError",
    "Control",
    "Mesh",
    "Model",
    "ModelBuilder",
    "NativeLibraryMissing",
    "SolverDiffPhD",
    "State",
]
