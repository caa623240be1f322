"""GPU: bench.py's output contract on a short run: one JSON line with the metric fields, the roofline
of the dominant kernel, the tuned tile shapes, the other forms' delta_p against FP64 and the CPU
baseline (sized down: a few timed steps, no sweep, no 8192^2 case or slab)."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_json_line():
    print("[gpu] bench.py --steps 20 --warmup 5 (f64 headline, f32z beside it)", flush=True)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--steps", "20", "--warmup", "5", "--forms", "f32z",
                        "--sweep-steps", "0", "--large-steps", "0", "--shard-steps", "0"],
                       cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [x for x in r.stdout.strip().splitlines() if x.strip()]
    assert len(lines) == 1, r.stdout[-2000:]
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
              "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 20 and d["warmup"] == 5
    assert d["value"] > 0 and d["ms_per_step"] > 0 and d["higher_is_better"] is True
    assert d["dtype"] == "f64" and d["config"]["storage"] == "f64"
    assert "workload" in d["config"] and "model" not in d["config"]
    roof = d["roofline"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in roof, k
    assert roof["bound"] == "hbm" and roof["unit"] == "GB/s" and roof["peak"] == 8000.0
    assert abs(roof["frac"] - roof["achieved"] / roof["peak"]) < 1e-9
    assert "frac_at_304B" not in roof
    assert 0 < roof["frac"] < 1 and roof["stream_GBs"] > 0
    cpu = d["cpu_baseline"]
    for k in ("value", "unit", "cores", "kind", "sample", "cpu_model", "nproc"):
        assert k in cpu, k
    assert cpu["kind"] in ("reference", "port") and cpu["value"] > 0
    assert set(d["forms"]) == {"f64", "f32z"}
    # the tuning steps ran before the timed ones, whatever --warmup is
    for form in d["forms"].values():
        assert form["tile_shapes"] and all(s["tuned"] for s in form["tile_shapes"].values()), form["tile_shapes"]
    # the launch form was chosen (FP64: by timing; FP32 / FP16: one launch per family)
    assert d["forms"]["f64"]["launch_form"] in (0, 1, 2) and d["forms"]["f32z"]["launch_form"] == 0
    # FP32 z-symmetric storage tracks FP64 in delta_p after the same ~426 steps (sharp drops, still relaxing)
    agree = d["forms"]["f32z"]["delta_p_vs_f64"]
    assert agree["cases"] == 14 and agree["max_rel"] < 1e-3, agree
