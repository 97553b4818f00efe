"""gpumatting_amd -- closed-form (Levin) matting on MI355X (gfx950).

The product is libgpumatting.so (HIP kernels behind a C ABI, include/*.h).
This package holds the ctypes binding (`_lib`), a thin device-array layer on
top of PyTorch tensors (`matting`), the seeded workload generator (`synth`)
and the autograd wrapper of the adjoint (`autograd`).
"""
from ._lib import (ASSEMBLY_ACCURATE, ASSEMBLY_REFERENCE, METHOD_CG, METHOD_GD, METHOD_PCG, PREC_BF16,
                   PREC_FP16, PREC_FP32, BandedMatrix, GMError, SolveInfo, SolveOpts, lib)

__all__ = ["lib", "GMError", "BandedMatrix", "SolveOpts", "SolveInfo", "METHOD_CG", "METHOD_PCG", "METHOD_GD",
           "PREC_FP32", "PREC_FP16", "PREC_BF16", "ASSEMBLY_REFERENCE", "ASSEMBLY_ACCURATE"]
