"""bench.py's contract on the CPU: one GPU per process, and the pieces of
the JSON line that do not need a GPU."""
import json
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU per process" in r.stderr
    assert r.stdout == ""  # no JSON line for an unsupported run


def test_kernel_name_follows_the_compiled_search_shape():
    class P:
        k, Ns, Npr, Nb = 7, 7, 5, 2
    assert bench.kernel_name(P, 3) == "k_bm<7, 3, 1>"
    P.Nb = 3
    assert bench.kernel_name(P, 1) == "k_bm<7, 1, 0>"


def test_traffic_table_matches_the_committed_profile():
    """The roofline's traffic comes from profiles/traffic.json for the
    default workload's dominant kernel; the summary beside it names the
    same kernel."""
    workload = ("60 RGB frames 1920x1080, sigma=20, both steps, default parameters, "
                "reference grid nb=1")
    t = bench.load_traffic(workload, "k_bm<7, 3, 1>")
    assert t is not None and t > 0
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
        ent = json.load(f)[workload]["k_bm<7, 3, 1>"]
    assert t == ent["fetch_bytes_per_launch"] + ent["write_bytes_per_launch"]
    with open(os.path.join(ROOT, "profiles", "vbm3d_1080p_rgb_sigma20.md")) as f:
        assert "`k_bm<7, 3, 1>`" in f.read()
    assert bench.load_traffic("no such workload", "k_bm<7, 3, 1>") is None
