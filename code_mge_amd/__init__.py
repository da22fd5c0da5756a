"""MI355X-native multigrid Jameson-RK solve step (code_mge mgsol path).

Host-side mirror of the reference's solver/euler/bcond call surface over
the C ABI in include/mge.h (libmge.so: HIP kernels for gfx950 plus the
cut-cell geometry preprocessor).
"""
from .params import Params
from .grid import LevelSpec, hierarchy, body, level_geometry
from .solver import Case, force_coefficients

__all__ = ["Params", "LevelSpec", "hierarchy", "body", "level_geometry", "Case", "force_coefficients"]
