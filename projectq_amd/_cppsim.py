"""Drop-in for `projectq.backends._sim._cppsim`: the `Simulator` class backed by the HIP library.

`Simulator(seed)` has the method surface of the reference's pybind11 class
(projectq/backends/_sim/_cppsim.cpp:43-67).  Extra keyword options select the MI355X features:

* precision="c128" (the reference's complex<double>) or "c64" (complex<float> state, FP64 sums);
* fusion="reference" (the reference's greedy min-4/max-5 rule, simulator.hpp:204-222) or
  "window" (gates queue until run(); a planner partitions the whole window);
* planner="greedy" (the reference's block rule, one pass per block), "tile" (LDS tile passes,
  many gates per pass) or "learned" (tile passes chosen with a learned cost model) -- window
  mode only;
* capacity_qubits=N reserves HBM for N qubits up front.

There is no CPU fallback: without the HIP library this module raises ImportError.
"""

import ctypes
import os

from ._binding import CAbiSimulator, bind_library

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libprojectq_amd.so")

_lib = None


def library():
    """Load libprojectq_amd.so (built by projectq_amd.build); raise ImportError if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m projectq_amd.build` (hipcc, gfx950)"
            )
        _lib = bind_library(ctypes.CDLL(LIB_PATH))
    return _lib


class Simulator(CAbiSimulator):
    """State-vector simulator resident in one GPU's HBM, with the `_cppsim.Simulator` surface."""

    def __init__(self, seed, precision="c128", fusion="reference", planner="greedy", capacity_qubits=0):
        super().__init__(library(), seed, precision, fusion, planner, capacity_qubits)
