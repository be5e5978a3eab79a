"""Loader for the C++ oracle (oracle/cpu_sim.cpp) -- test infrastructure only."""

import ctypes
import os
import subprocess

from projectq_amd._binding import CAbiBatch, CAbiSimulator, bind_library

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_sim.so")

_lib = None


def build(force=False):
    """Compile oracle/cpu_sim.cpp with g++ (make -C oracle)."""
    cmd = ["make", "-C", HERE] + (["-B"] if force else []) + ["liboracle_sim.so"]
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)


def library():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = bind_library(ctypes.CDLL(LIB_PATH))
    return _lib


class OracleSimulator(CAbiSimulator):
    """The reference engine restated on the CPU, behind the `_cppsim.Simulator` surface."""

    def __init__(self, seed):
        super().__init__(library(), seed, precision="c128", fusion="reference", planner="greedy")


class OracleBatch(CAbiBatch):
    """Batched circuits on the CPU oracle: each circuit unfused, then its expectation value."""

    def __init__(self, circuits, operators):
        super().__init__(library(), circuits, operators, precision="c128", autotune=False)
