"""am_ml_suite_amd — MI355X-native SIMP topology-optimization engine (libamt.so).

The hot path (matrix-free hex8 operator, Jacobi-PCG, sensitivity filter, OC bisection) runs
as hand-written gfx950 HIP kernels behind the C-ABI in include/amt.h. This package is the
Python mirror of the reference optimizer API over that ABI (ctypes).
"""
from .optimizer import (  # noqa: F401
    AmtError, BatchOptimizer, CPPTopologyOptimizer, IterationResult, MaterialProperties, OptimizationResult,
    SIMPConfig, TopologyOptimizer, UNet, device_name, element_stiffness, get_max_threads,
    is_available)
from .problems import (  # noqa: F401
    LoadCase, VoxelDomain, bracket_domain, bracket_load_case, cantilever_problem,
    bracket_like_problem, batch_problems)
