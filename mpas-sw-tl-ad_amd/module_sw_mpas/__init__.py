"""Drop-in replacement for the reference's f2py extension ``module_sw_mpas``.

``from module_sw_mpas import mpas_sw_module as mpsw`` gives an object with
the reference module's attribute interface (``src/mpas_sw_module.f90``) whose
stepping routines run on the GPU (see ``mpas_sw_amd``).
"""

from ._facade import mpas_sw_module  # noqa: F401
