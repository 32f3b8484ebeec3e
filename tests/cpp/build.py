"""Builds tests/cpp/dropin_driver: a C++ program that uses the drop-in classes through
include/*.h and libasd2_hip.so only, as a reference driver would."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "dropin_driver.cpp")
EXE = os.path.join(HERE, "dropin_driver")
LIBDIR = os.path.join(ROOT, "asd2_amd")


def build(verbose: bool = False) -> str:
    cxx = shutil.which("g++") or "g++"
    cmd = [cxx, "-O2", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}", SRC, "-o", EXE + ".tmp",
           f"-L{LIBDIR}", "-lasd2_hip", "-Wl,-rpath,$ORIGIN/../../asd2_amd"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("dropin_driver build failed")
    os.replace(EXE + ".tmp", EXE)
    return EXE


if __name__ == "__main__":
    print(build(verbose=True))
