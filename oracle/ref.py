"""Loader for ProjectQ's own C++ simulator compiled by `make -C oracle ref` -- test infrastructure.

The module is the reference's pybind11 `_cppsim` (projectq/backends/_sim/_cppsim.cpp) built from
the reference sources into oracle/_ref/<variant>/.  Returns None where that build is absent.
"""

import importlib.util
import glob
import os

HERE = os.path.dirname(os.path.abspath(__file__))
_cache = {}


def available(variant="intrin"):
    """Whether the build exists, without loading it."""
    return bool(glob.glob(os.path.join(HERE, "_ref", variant, "_cppsim*.so")))


def load(variant="intrin"):
    if variant in _cache:
        return _cache[variant]
    mod = None
    paths = glob.glob(os.path.join(HERE, "_ref", variant, "_cppsim*.so"))
    if paths:
        try:
            spec = importlib.util.spec_from_file_location("_cppsim", paths[0])
            mod = importlib.util.module_from_spec(spec)
            spec.loader.exec_module(mod)
        except ImportError:
            mod = None
    _cache[variant] = mod
    return mod
