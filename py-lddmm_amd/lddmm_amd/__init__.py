"""lddmm_amd: MI355X (gfx950) HIP implementation of py-lddmm's point-set hot path.

Module names mirror the reference package (py-lddmm/base): kernelFunctions,
kernelFunctions_util, surfaces, curves, surfaceDistances, curveDistances,
pointEvolution, bfgs, linesearch, surfaceMatching, curveMatching,
secondOrderPointSetMatching.  All kernel sums run through liblddmm_amd.so.
"""
from . import _backend  # noqa: F401  (fails loudly when the HIP library is missing)

__version__ = "0.1.0"
