"""dioptrin_amd: MI355X-native azimuthal-integration hot path.

Device-built CSR lookup tables and HIP gather kernels for gfx950 behind a
C ABI (include/dioptrin_hip.h), with a Python front end that mirrors the
reference Integrator / GPU-plan API.
"""
from ._native import (DioError, DioUnavailableError, DioUnsupportedError, LIB_PATH)
from .integrator import (BatchResult, GpuGraph, GpuPlan, IntegrationResult, Integrator,
                         device_info, load_poni)

__all__ = ["Integrator", "GpuPlan", "GpuGraph", "IntegrationResult", "BatchResult",
           "device_info", "load_poni", "DioError", "DioUnavailableError",
           "DioUnsupportedError", "LIB_PATH"]
