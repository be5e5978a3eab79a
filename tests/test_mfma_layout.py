"""CPU check of the MFMA dense-block layout (projectq_amd/csrc/mfma_layout.hpp): the per-lane matrix
table, the kernel's per-lane loads and stores ((re, im) pairs of half a group per lane), and the
documented lane semantics of v_mfma_f32_32x32x2f32 (k = 3, 4, 5),
emulated in FP64 (tests/native/mfma_emu.cpp), must give M v for every group.  A row-order or
sign mistake in the table fails here; the GPU tests then check the hardware against the oracle."""

import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NATIVE = os.path.join(ROOT, "tests", "native")
LIB = os.path.join(NATIVE, "_build", "libmfma_emu.so")


@pytest.fixture(scope="module")
def emu():
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", os.path.join(ROOT, "projectq_amd", "csrc"),
                    os.path.join(NATIVE, "mfma_emu.cpp"), "-o", LIB], check=True)
    return ctypes.CDLL(LIB)


@pytest.mark.parametrize("k", [3, 4, 5])
@pytest.mark.parametrize("n_groups", [1, 37, 70])
def test_mfma_layout_applies_the_block(emu, k, n_groups):
    rng = np.random.default_rng(10 * k + n_groups)
    d = 2**k
    m = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))  # not unitary: every entry matters
    v = rng.normal(size=(n_groups, d)) + 1j * rng.normal(size=(n_groups, d))
    want = v @ m.T  # row g: M v_g
    buf = np.ascontiguousarray(v.astype(np.complex128))
    mm = np.ascontiguousarray(m.astype(np.complex128))
    emu.mfma_emulate(ctypes.c_int(k), mm.ctypes.data_as(ctypes.c_void_p), buf.ctypes.data_as(ctypes.c_void_p),
                     ctypes.c_int(n_groups))
    # the table holds FP32 entries: compare at FP32 accuracy
    assert np.max(np.abs(buf - want)) <= 1e-5 * np.max(np.abs(want))
