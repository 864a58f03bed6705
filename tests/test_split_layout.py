"""The row-block layout behind pdlp mode's split rows, checked on the CPU:
tests/layout/split_layout_check.hip (host code against
nirnaya_amd/csrc/engine/devmat.hpp's segment_blocks) is compiled with hipcc
and run; it exits 0 only when every piece covers its nonzeros once, starts on
a group boundary of its row, fits max(tile, 2,048) nonzeros and writes only
its row's group slots."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_split_layout_invariants(tmp_path):
    exe = tmp_path / "split_layout_check"
    src = os.path.join(HERE, "layout", "split_layout_check.hip")
    subprocess.run([HIPCC, "-std=c++17", "-O1", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", src, "-o", str(exe)], check=True, capture_output=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 violations" in r.stdout
