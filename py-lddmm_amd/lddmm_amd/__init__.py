"""lddmm_amd: MI355X (gfx950) HIP implementation of py-lddmm's point-set hot path.

Module names mirror the reference package (py-lddmm/base): kernelFunctions,
kernelFunctions_util, surfaces, curves, surfaceDistances, curveDistances,
pointEvolution, bfgs, linesearch, surfaceMatching, curveMatching,
secondOrderPointSetMatching.  All kernel sums run through liblddmm_amd.so (loaded by _backend on first
use of a kernel module; a missing library raises ImportError).
"""
__version__ = "0.1.0"
