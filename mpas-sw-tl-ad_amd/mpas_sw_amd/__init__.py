"""mpas_sw_amd: MI355X (gfx950) MPAS shallow-water RK4 core with its
tangent-linear model and adjoint.

Host side: mesh generation and space-filling-curve ordering (``mesh``),
Williamson initial conditions (``testcases``), reconstruction set-up
(``reconstruct``); device side: ``SWModel`` over the HIP library
(``core``/``_lib``).
"""

from .mesh import Mesh, icosahedral_mesh, get_mesh, sfc_permutation  # noqa: F401
from .testcases import initial_conditions  # noqa: F401
from .reconstruct import reconstruct_coefficients  # noqa: F401


def __getattr__(name):
    # the torch/HIP side is imported lazily so that host-only use (mesh,
    # initial conditions) does not need the built library
    if name == "SWModel":
        from .core import SWModel
        return SWModel
    raise AttributeError(name)
