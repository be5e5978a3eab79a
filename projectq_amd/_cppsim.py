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


MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "planner_mlp.npz")
_model = None


def planner_model():
    """Weights of the learned planner's cost model (projectq_amd/planner_mlp.npz), or None."""
    global _model
    if _model is None and os.path.exists(MODEL_PATH):
        import numpy as np

        with np.load(MODEL_PATH, allow_pickle=False) as f:
            _model = np.asarray(f["weights"], dtype=np.float32)
    return _model


class Simulator(CAbiSimulator):
    """State-vector simulator resident in one GPU's HBM, with the `_cppsim.Simulator` surface."""

    def __init__(self, seed, precision="c128", fusion="reference", planner="greedy", capacity_qubits=0,
                 planner_weights=None):
        super().__init__(library(), seed, precision, fusion, planner, capacity_qubits)
        if planner == "learned":
            w = planner_weights if planner_weights is not None else planner_model()
            if w is not None:
                self.set_planner_model(w)
