"""MI355X-native (gfx950) shallow-water evolve() for ANUGA-style domains.

A drop-in for the step path of the reference GPU domain
(src/anuga_cuda/gpu_domain_advanced.py): ANUGA's Domain API on top of a
C ABI (include/anuga_sw.h) whose every stage is a hand-written HIP kernel.
"""
from .domain import (DEFAULT_PARAMETERS, Dirichlet_boundary, Domain, Quantity,
                     Reflective_boundary, Transmissive_boundary, create_domain_from_file)
from .ensemble import Ensemble
from .mesh import Mesh, delaunay_random, read_mesh, read_msh, read_tsh, rectangular_cross

GPU_domain = Domain

__all__ = [
    "Domain", "GPU_domain", "Quantity", "Mesh", "Reflective_boundary", "Dirichlet_boundary",
    "Transmissive_boundary", "rectangular_cross", "read_tsh", "read_msh", "read_mesh", "delaunay_random",
    "DEFAULT_PARAMETERS", "Ensemble", "create_domain_from_file",
]
