"""bench.py's output contract: one JSON line with the required keys, the
roofline and CPU-baseline objects, and a refusal of more than one GPU.
The refusal runs on the CPU; the short default run needs the GPU."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


@pytest.mark.gpu
def test_default_run_prints_the_contract_line():
    # A short run of the default workload (the 400,000 x 1,200,000 LP) with
    # the CPU baseline's bounded sample, without the time-to-tolerance solves.
    r = subprocess.run([sys.executable, "-u", BENCH, "--steps", "50", "--warmup", "10", "--no-ttt", "--no-batched"],
                       capture_output=True, text=True, timeout=600, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    assert len(lines) == 1, r.stdout[-2000:]
    d = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
                "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert key in d, key
    assert d["n_gpus"] == 1 and d["steps"] == 50 and d["warmup"] == 10
    assert d["value"] > 0 and d["higher_is_better"] is True
    assert abs(d["value"] * d["ms_per_step"] / 1e3 - 1.0) < 0.01   # iterations/s and ms per iteration agree
    assert "workload" in d["config"]
    rf = d["roofline"]
    for key in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert key in rf, key
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-3
    cb = d["cpu_baseline"]
    for key in ("value", "unit", "cores", "kind", "sample"):
        assert key in cb, key
    assert cb["kind"] in ("reference", "port") and cb["value"] > 0
    # The longer production sweep beside the timed steps agrees with `value`
    # to within run-to-run noise.
    ps = d["production_sweep"]
    assert ps["iterations"] == 2000 and abs(ps["iterations_per_s"] / d["value"] - 1.0) < 0.2


@pytest.mark.gpu
@pytest.mark.parametrize("argv", [
    ["--workload", "mixed", "--steps", "16", "--warmup", "4", "--no-ttt", "--no-cpu-baseline"],
    ["--workload", "family", "--steps", "64", "--fam-lanes", "128", "--ws-m", "5000", "--ws-n", "12000"],
    ["--workload", "dense-row", "--steps", "64", "--warmup", "8", "--dense-len", "0,20000"],
    ["--workload", "pdlp-fp64", "--steps", "64", "--warmup", "8"],
], ids=["mixed", "family", "dense-row", "pdlp-fp64"])
def test_other_workloads_print_one_line(argv):
    # The secondary workloads' lines (mixed precision on the large LP, the
    # family engine on a small family, the dense-row LPs) parse and carry a
    # positive value.
    r = subprocess.run([sys.executable, "-u", BENCH] + argv, capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    assert len(lines) == 1, r.stdout[-2000:]
    d = json.loads(lines[0])
    assert d["value"] > 0 and d["n_gpus"] == 1 and "workload" in d["config"]
