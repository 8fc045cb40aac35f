"""dioptrin_amd: MI355X-native azimuthal-integration hot path.

Device-built CSR lookup tables and HIP gather kernels for gfx950 behind a
C ABI (include/dioptrin_hip.h), with a Python front end that mirrors the
reference Integrator / GPU-plan API.
"""
from ._native import (DioError, DioUnavailableError, DioUnsupportedError, LIB_PATH)
from .grid import (IntegrationGrid2d, IntermediateResult1d, IntermediateResult2d,
                   azimuthal_bracket_bins)
from .integrator import (BatchResult, GpuGraph, GpuPlan, IntegrationResult, Integrator,
                         device_info, load_poni)
from .multi import MultiIntegrator

__all__ = ["Integrator", "MultiIntegrator", "IntegrationGrid2d", "IntermediateResult1d",
           "IntermediateResult2d", "azimuthal_bracket_bins", "GpuPlan", "GpuGraph",
           "IntegrationResult", "BatchResult",
           "device_info", "load_poni", "DioError", "DioUnavailableError",
           "DioUnsupportedError", "LIB_PATH"]
