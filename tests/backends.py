"""Backend factories for the parametrised suite (mirrors the `sim` fixture of
projectq/backends/_sim/_simulator_test.py:80-93, which swaps `Simulator._simulator`)."""

import pytest

CPU_BACKENDS = ["oracle"]
GPU_BACKENDS = ["hip_c128", "hip_c64"]

# tolerance each backend's amplitudes meet against exact answers
TOL = {"oracle": 1e-12, "hip_c128": 1e-12, "hip_c64": 2e-6}


def make(kind, seed=1, **kw):
    if kind == "oracle":
        from oracle.cpu_sim import OracleSimulator

        return OracleSimulator(seed)
    import projectq_amd

    if kind == "hip_c128":
        return projectq_amd.Simulator(seed, precision="c128", **kw)
    if kind == "hip_c64":
        return projectq_amd.Simulator(seed, precision="c64", **kw)
    raise ValueError(kind)


def params(include_gpu=True):
    out = [pytest.param(k, id=k) for k in CPU_BACKENDS]
    if include_gpu:
        out += [pytest.param(k, id=k, marks=pytest.mark.gpu) for k in GPU_BACKENDS]
    return out
