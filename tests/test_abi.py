"""The C-ABI libraries load and export every symbol include/projectq_amd.h declares (CPU only:
no compute calls are made)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "projectq_amd.h")).read()
    return sorted(set(re.findall(r"\b(qs_[a-z_A-Z]+)\s*\(", text)) - {"qs_math_fn"})


def test_header_matches_binding():
    from projectq_amd._binding import ABI_SYMBOLS

    assert sorted(ABI_SYMBOLS) == declared_symbols()


def test_hip_library_exports_every_symbol():
    import projectq_amd

    lib = projectq_amd.library()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_oracle_library_exports_every_symbol():
    from oracle import cpu_sim

    lib = cpu_sim.library()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_hip_library_is_gfx950():
    path = os.path.join(ROOT, "projectq_amd", "libprojectq_amd.so")
    data = open(path, "rb").read()
    assert b"gfx950" in data


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    import projectq_amd._cppsim as cs

    monkeypatch.setattr(cs, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(cs, "_lib", None)
    with pytest.raises(ImportError):
        cs.library()


def test_abi_version():
    import projectq_amd

    assert projectq_amd.library().qs_abi_version() == 1
    assert ctypes.c_int(1).value == 1
