"""b12_amd — MI355X-native implementation of B12's set-oriented hot path.

The compute runs in the HIP library libb12x.so (built by b12_amd.build); this package only
binds its C ABI (include/b12x_c_api.h) for Python callers, tests and the benchmark.
"""
from .api import (Forest, serve_run, continuation, chainRecurrentSet, eigs, KrylovBasis, fixedBoxes, subdivisionByNim, ALL, EQU, EXPD, EXTR, ENTR, HIT, INS, INVALID, LEAF, NONE, SD, B12xError, CooMatrix, Duffing,
                  Henon, HenonHeiles, Identity, ImplicitBoxTree, InnerGrid, Lorenz, PointIterator, Scheme,
                  UniformGrid, bench_copy, compressIndexVector, device_name, library_path, load,
                  power_iteration_form, power_iteration_kernel_ms, tm_last_passes)

__all__ = ["Forest", "serve_run", "continuation", "chainRecurrentSet", "eigs", "KrylovBasis", "fixedBoxes", "subdivisionByNim", "ALL", "EQU", "EXPD", "EXTR", "ENTR", "HIT", "INS", "INVALID", "LEAF", "NONE", "SD", "B12xError",
           "CooMatrix", "Duffing", "Henon", "HenonHeiles", "Identity", "ImplicitBoxTree", "InnerGrid", "Lorenz",
           "PointIterator", "Scheme", "UniformGrid", "bench_copy", "compressIndexVector", "device_name", "library_path", "load",
           "power_iteration_form", "power_iteration_kernel_ms", "tm_last_passes"]
