"""bench.py's command-line contract on the host: more than one GPU is refused with a
message (this build measures one MI355X), and the PMC traffic summaries it reads belong
to the headline lattice and are internally consistent."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


@pytest.mark.parametrize("dtype,bpu", [("f32", 152), ("f16", 76)])
def test_traffic_summaries(dtype, bpu):
    p = os.path.join(ROOT, "profiles", f"k_step_porous_wall_640x320x160_{dtype}_traffic.json")
    with open(p) as f:
        t = json.load(f)
    cells = 640 * 320 * 160
    assert t["algorithmic_bytes"] == bpu * cells
    assert t["bytes_per_launch"] == pytest.approx(t["fetch_bytes"] + t["write_bytes"], abs=2)
    assert t["fetch_bytes"] == pytest.approx(t["fetch_size_kib_raw"] * 1024 * t["fetch_scale"], rel=1e-4)
    assert t["write_bytes"] == pytest.approx(t["write_size_kib_raw"] * 1024 * t["write_scale"], rel=1e-4)
    # measured traffic lies within a few per cent of the algorithmic bytes
    assert 0.95 < t["bytes_per_launch"] / t["algorithmic_bytes"] < 1.10
    assert all("bench.py" in c for c in t["commands"])


def test_two_step_traffic_summary():
    """k_step2 advances two steps per launch: its HBM traffic lies between the bytes of one
    pass over the populations (152 B per cell) and those of two single steps."""
    p = os.path.join(ROOT, "profiles", "k_step2_porous_wall_640x320x160_f32_traffic.json")
    with open(p) as f:
        t = json.load(f)
    cells = 640 * 320 * 160
    assert t["steps_per_launch"] == 2
    assert t["algorithmic_bytes"] == 2 * 152 * cells
    assert t["bytes_per_launch"] == pytest.approx(t["fetch_bytes"] + t["write_bytes"], abs=2)
    assert t["fetch_bytes"] == pytest.approx(t["fetch_size_kib_raw"] * 1024 * t["fetch_scale"], rel=1e-4)
    assert t["write_bytes"] == pytest.approx(t["write_size_kib_raw"] * 1024 * t["write_scale"], rel=1e-4)
    assert 152 * cells < t["bytes_per_launch"] < t["algorithmic_bytes"]
    # every population is written exactly once per launch
    assert t["write_bytes"] == pytest.approx(76 * cells, rel=0.02)
    assert all("bench.py" in c for c in t["commands"])
