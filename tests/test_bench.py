"""bench.py's output contract: one JSON line with the driver's keys, a
roofline object and a cpu_baseline object; one GPU only."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KEYS = {"metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
        "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config"}


def run(args, timeout=600):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + args, cwd=ROOT,
                          capture_output=True, text=True, timeout=timeout)


def test_more_than_one_gpu_is_refused():
    r = run(["--gpus", "2"], timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


@pytest.mark.gpu
def test_batch_line_contract():
    r = run(["--count", "48", "--cap-epochs", "4", "--no-tune", "--cpu-seconds", "1",
             "--plain-count", "8"])
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    assert KEYS <= set(d) and d["n_gpus"] == 1 and d["unit"] == "matrices/s"
    assert d["config"]["workload"].startswith("batch")
    rf = d["roofline"]
    assert {"bound", "achieved", "peak", "unit", "frac", "traffic"} <= set(rf)
    assert rf["frac"] == pytest.approx(rf["achieved"] / rf["peak"])
    cb = d["cpu_baseline"]
    if cb is not None:  # the reference binary is built only where its sources are
        assert {"value", "unit", "cores", "kind", "sample"} <= set(cb)
        assert len(cb["mode_i"]["runs"]) == len(cb["mode_ii"]["runs"]) == 3
    assert d["verified"] == d["solved"]
    pf = d["plain_form"]
    assert pf["instances"] == 8 and d["batch_over_plain"] == pytest.approx(
        d["value"] / max(pf["value"], 1e-12))
    assert sum(b["instances"] for b in d["by_order"].values()) == 48


@pytest.mark.gpu
def test_dual_line_contract():
    r = run(["--workload", "dual", "--seconds", "1", "--no-cpu"])
    assert r.returncode == 0, r.stderr[-2000:]
    d = json.loads(r.stdout.strip().splitlines()[-1])
    assert KEYS <= set(d) and d["unit"] == "moves/s" and d["verify_status"] == 0
    assert d["roofline"]["kernel"] == "k_anneal_dual"
