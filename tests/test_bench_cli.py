"""bench.py's host-side contract (no GPU): one GPU only, the measured-traffic
lookup behind roofline.traffic, the records' sources, the step table and the
algorithmic bytes per node-step."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    proc = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                          capture_output=True, text=True, timeout=120)
    assert proc.returncode == 2
    assert "one GPU" in proc.stderr and proc.stdout == ""


def test_traffic_records_name_existing_profiles():
    with open(bench.TRAFFIC_FILE) as f:
        records = json.load(f)["records"]
    assert records
    for r in records:
        assert os.path.exists(os.path.join(REPO, r["source"])), r["source"]
        per_node = r.get("bytes_per_node_launch", r.get("bytes_per_node_step"))
        # the measured bytes are the sum of the read and written parts; a
        # reaction launch reads and writes every state once, which no cache
        # holds between launches; a fused launch keeps Vm in LDS and streams
        # the gates through the L2 caches, so its HBM bytes per node-step may
        # fall well below the split step's
        assert abs(per_node - (r["read"] + r["write"])) < 0.1, r
        if r["kernel_class"] == "fused":
            assert 0 < per_node <= 2 * r["algorithmic"], r
        else:
            assert per_node >= r["algorithmic"], r


def test_measured_traffic_scales_per_launch_and_per_node_step(tmp_path, monkeypatch):
    f = tmp_path / "traffic.json"
    f.write_text(json.dumps({"records": [
        {"workload": "w", "kernel_class": "reaction_pair", "threads": 64, "bytes_per_node_launch": 900.0,
         "source": "s1"},
        {"workload": "w", "kernel_class": "fused", "threads": 256, "bytes_per_node_step": 40.0, "source": "s2"}]}))
    monkeypatch.setattr(bench, "TRAFFIC_FILE", str(f))
    assert bench.measured_traffic("w", "reaction_pair", 64, 1000, 5e5) == (900.0 * 1000, "s1")
    assert bench.measured_traffic("w", "fused", 256, 1000, 512 * 1000) == (40.0 * 512 * 1000, "s2")
    assert bench.measured_traffic("w", "reaction_pair", 256, 1000, 0) == (None, None)
    assert bench.measured_traffic("other", "fused", 256, 1000, 0) == (None, None)
    monkeypatch.setattr(bench, "TRAFFIC_FILE", str(tmp_path / "missing.json"))
    assert bench.measured_traffic("w", "fused", 256, 1000, 1) == (None, None)


def test_step_table_and_algorithmic_bytes():
    t = bench.step_table(0.5, 2.5e-5, 4)
    assert t.shape == (4, 4)
    np.testing.assert_array_equal(t[:, 0], 0.5 + np.arange(4) * 2.5e-5)
    assert np.all(t[:, 1] == 2.5e-5) and np.all(t[:, 2] == 1.25e-5) and np.all(t[:, 3] == 2.5e-5)
    # every state read and written once: FP64 16 B per state; mixed 8 B per gate + 16 B Vm
    assert bench.bytes_per_node_step(56, "fp64") == 896
    assert bench.bytes_per_node_step(56, "mixed") == 456
    assert bench.bytes_per_node_step(4, "fp64") == 64


@pytest.mark.parametrize("kernel_file", ["k_reaction_fibre_pair_shorten56_1024x1481.md",
                                         "k_reaction_fibre_pair_tk2014_1024x1481.md",
                                         "k_reaction_pair_and_k_fused_hh_4096_uneven.md",
                                         "k_adj_fused_hh_256x1000.md"])
def test_profile_summaries_give_their_commands(kernel_file):
    """each summary names the rocprofv3 commands it comes from, and those run
    only committed files (bench.py)"""
    with open(os.path.join(REPO, "profiles", kernel_file)) as f:
        text = f.read()
    cmds = [l.strip() for l in text.splitlines() if l.strip().startswith("rocprofv3 ")]
    assert cmds, kernel_file
    for c in cmds:
        prog = c.split(" -- ", 1)[1].split()
        assert prog[0] == "python3" and prog[1] == "bench.py", c
        assert os.path.exists(os.path.join(REPO, prog[1]))
