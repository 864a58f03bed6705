"""NIRNAYA-MI355X: a GPU PDHG engine for LPs behind NIRNAYA's C ABI.

Native code lives in nirnaya_amd/csrc (host C++17 + HIP kernels for gfx950)
and is built into bin/libsov.so. This package holds the Python mirrors of the
reference's interfaces:
  nirnaya_amd.nirnaya  NirnayaModel (python/nirnaya.py of the reference)
  nirnaya_amd.engine   direct binding of the engine C ABI (include/nirnaya/pdhg.h)
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "nirnaya_amd", "csrc")


def build(jobs: int = 8) -> None:
    """Compile libsov.so, sovsolve and checker for gfx950 (make in csrc)."""
    subprocess.run(["make", "-C", CSRC, f"-j{jobs}"], check=True)
