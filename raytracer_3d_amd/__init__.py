"""MI355X-native (gfx950, HIP) 6D phase-space wave-energy time step.

Drop-in for the hot path of the reference Fortran model: calc_parameter,
integrate/superbee_fluxes and the diagnose reductions, behind the C ABI of
include/rt6d.h.  ``rt6d`` holds the Python host bindings, ``cases`` the
synthetic integrations, ``fortran/`` the Fortran host bindings and driver.
"""
from . import cases  # noqa: F401

__all__ = ["cases", "rt6d"]
