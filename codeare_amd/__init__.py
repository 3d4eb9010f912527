"""codeare_amd -- MI355X (gfx950) HIP implementation of codeare's non-Cartesian
CG-SENSE path: Kaiser-Bessel NFFT forward/adjoint, the NCSENSE coil encoding
operator E / E^H and its CG solve, behind the C ABI in include/cgsense.h.
"""
from .cgsense import (NFFT, NCSENSE, Plan, CgsError, load_library, LIB_PATH,  # noqa: F401
                      CGS_EH_TRUE, CGS_EH_REF, CGS_TOL_NORM_RATIO, CGS_TOL_SQUARED_RATIO,
                      CGS_OK, CGS_ERR_ARG, CGS_ERR_STATE, CGS_ERR_HIP, CGS_ERR_ALLOC, CGS_ERR_UNSUPPORTED,
                      CGS_ERR_NAN)

__all__ = ["NFFT", "NCSENSE", "Plan", "CgsError", "load_library", "LIB_PATH",
           "CGS_EH_TRUE", "CGS_EH_REF", "CGS_TOL_NORM_RATIO", "CGS_TOL_SQUARED_RATIO",
           "CGS_OK", "CGS_ERR_ARG", "CGS_ERR_STATE", "CGS_ERR_HIP", "CGS_ERR_ALLOC", "CGS_ERR_UNSUPPORTED",
           "CGS_ERR_NAN"]
