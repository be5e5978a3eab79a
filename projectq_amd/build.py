"""Build recipe of the HIP library (gfx950 only): `python -m projectq_amd.build`.

Compiles projectq_amd/csrc/* with hipcc into projectq_amd/libprojectq_amd.so, in tree, so the
library travels with the source tree.  Objects go to projectq_amd/_build/.
"""

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libprojectq_amd.so")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fvisibility=hidden",
    # packed FP32 VALU ops (v_pk_*_f32) stay off: ordinary float code only
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    # uniform branches (the register-op dispatch) keep their CFG: no flow blocks whose phis copy
    # the amplitude registers
    "-mllvm",
    "-structurizecfg-skip-uniform-regions=true",
    "-I",
    os.path.join(ROOT, "include"),
    "-I",
    CSRC,
    "-Wall",
    "-Wno-unused-command-line-argument",
]

SOURCES = ["kernels.hip", "mlp.hip", "compare.hip", "noalloc.hip", "mfma.hip", "batch_energy.hip", "engine.cpp", "planner.cpp", "batch.cpp", "capi.cpp"]


def _compile(src):
    obj = os.path.join(BUILD, src + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{proc.stderr[-8000:]}")
    return obj


def build(verbose=False):
    """Compile every HIP/C++ source for gfx950 and link libprojectq_amd.so (always rebuilds)."""
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(len(SOURCES), 8)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stderr[-8000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
