"""MI355X (gfx950) engine for BioSpring's per-step force-and-integrate path.

The engine is libbiospring_amd.so (HIP kernels + C ABI, include/bs_api.h);
this package holds its Python binding (_abi), a SpringNetwork mirror over the
ABI (network) and the seeded workload generator (workloads).
"""
__version__ = "0.1.0"
