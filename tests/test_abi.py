"""The C-ABI library loads on a CPU-only host and exports every entry point declared in
include/*.h (no device calls are made here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "asd2_amd", "libasd2_hip.so")
HEADER = os.path.join(ROOT, "include", "asd2_hip.h")


def declared_c_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(asd2_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_c_symbols()
    for s in ("asd2_create", "asd2_ls_advect", "asd2_ls_reinit_minrk2", "asd2_ls_reinit_sussman",
              "asd2_gfm_assemble", "asd2_poisson_solve", "asd2_poisson_bicgstab", "asd2_csr_pcg",
              "asd2_mac_update_vel", "asd2_mac_intermediate_vel"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build the library first (python -m asd2_amd.build)"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [s for s in declared_c_symbols() if s not in exported]
    assert not missing, missing


def test_library_loads_without_gpu():
    lib = ctypes.CDLL(LIB)
    lib.asd2_version.restype = ctypes.c_int
    assert lib.asd2_version() >= 1
    lib.asd2_ls_reinit_iterations.argtypes = [ctypes.c_double, ctypes.c_double]
    lib.asd2_ls_reinit_iterations.restype = ctypes.c_int
    assert lib.asd2_ls_reinit_iterations(1.0 / 128, 2.25 / 128) == 5


def test_python_binding_fails_loudly_without_library(tmp_path, monkeypatch):
    import asd2_amd
    monkeypatch.setattr(asd2_amd, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(asd2_amd, "_lib", None)
    with pytest.raises(FileNotFoundError):
        asd2_amd.load()
