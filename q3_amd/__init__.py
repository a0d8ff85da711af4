"""q3_amd -- MI355X-native H|psi> and block-Davidson path of camsai/q3.

The compute path is libq3x.so (hand-written gfx950 HIP kernels + host C++17, C ABI in
include/q3x.h).  Importing this package does not touch the GPU; the library is loaded on
first use and its absence raises ImportError (there is no CPU fallback).
"""
from .planewave import (FP64, MIXED, BandRequestSpec, BandStructureServer, Context, Crystal,  # noqa: F401
                        KPoint, KPointBatch, PlanewaveBasis, PlanewaveDavidsonDiagonalizer,
                        PlanewaveHamiltonianApplicator)

__all__ = ["FP64", "MIXED", "Context", "Crystal", "KPoint", "PlanewaveBasis", "PlanewaveHamiltonianApplicator",
           "PlanewaveDavidsonDiagonalizer", "KPointBatch",
           "BandRequestSpec", "BandStructureServer"]
