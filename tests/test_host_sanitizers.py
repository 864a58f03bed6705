"""The host shell under AddressSanitizer and UndefinedBehaviorSanitizer: the
host sources the checker and the facade share (MPS and .nlpb I/O,
validation, presolve/postsolve, geometric scaling, the double-double
checker) are compiled with g++ -fsanitize=address,undefined (no HIP code
involved) into tests/sanitize/host_checks and into the checker, and run on
the crude-blending fixture and on 200 seeded random models. A sanitizer
report aborts the program (-fno-sanitize-recover=all), which fails the test.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "nirnaya_amd", "csrc", "host")
SOURCES = ["core", "model", "io", "lpb", "scaling", "presolve"]
FLAGS = ["-std=c++17", "-O1", "-g", "-ffp-contract=off", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-I" + HOST]
FIXTURE = os.path.join(ROOT, "tests", "golden", "MRPL_Crude_Blending_LP.mps")
FIXTURE_SOL = os.path.join(ROOT, "tests", "golden", "MRPL_Crude_Blending_LP.sol")


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("san")
    procs = [subprocess.Popen(["g++", *FLAGS, "-c", os.path.join(HOST, s + ".cpp"), "-o", str(d / (s + ".o"))])
             for s in SOURCES]
    assert all(p.wait() == 0 for p in procs)
    objs = [str(d / (s + ".o")) for s in SOURCES]
    subprocess.run(["g++", *FLAGS, os.path.join(ROOT, "tests", "sanitize", "host_checks.cpp"), *objs,
                    "-o", str(d / "host_checks")], check=True)
    subprocess.run(["g++", *FLAGS, os.path.join(HOST, "checker_main.cpp"), *objs, "-o", str(d / "checker")],
                   check=True)
    return d


def _env():
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    return env


def test_host_shell_clean_under_sanitizers(built, tmp_path):
    r = subprocess.run([str(built / "host_checks"), FIXTURE, str(tmp_path)], capture_output=True, text=True,
                       env=_env(), timeout=300)
    sys.stdout.write(r.stdout + r.stderr[-4000:])
    assert r.returncode == 0
    assert "0 failure(s)" in r.stdout
    n = int(r.stdout.split("host checks: ")[1].split()[0])
    assert n >= 100   # most random models pass validation and run every step


def test_checker_clean_under_sanitizers(built):
    r = subprocess.run([str(built / "checker"), FIXTURE, FIXTURE_SOL], capture_output=True, text=True, env=_env(),
                       timeout=120)
    sys.stdout.write(r.stdout + r.stderr[-4000:])
    assert r.returncode == 0 and "PASS" in r.stdout
