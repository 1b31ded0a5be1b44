"""MI355X-native implicit Newton-Raphson step for growing prism tissues (TissueFolding's hot path).

Product path: ``libtfnr.so`` (hand-written HIP for gfx950 behind the C ABI in include/tf_nr.h),
driven from C++ (include/tf_nr.hpp) or from Python through :class:`NewtonRaphsonSolver`.
"""
from .configs import Tissue, config1, config2, disc_tissue, wing_disc_tissue  # noqa: F401
from .mesh import TissueMesh, disc_mesh, ellipse_mesh, read_mesh, write_mesh  # noqa: F401


def NewtonRaphsonSolver(*args, **kw):  # noqa: N802 - mirrors the reference class name
    from .solver import NewtonRaphsonSolver as _S
    return _S(*args, **kw)
