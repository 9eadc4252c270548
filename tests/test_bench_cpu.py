"""bench.py's contract checked without a GPU: more than one GPU is refused with
a status and a message before anything touches a device, and the roofline's
`traffic` comes from the committed HBM counter summary (FETCH_SIZE and
WRITE_SIZE per dispatch of the two-sweep kernel at 256^3), above the bytes a
launch must move."""
import json
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, stdin=subprocess.DEVNULL)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_roofline_traffic_from_the_counter_summary():
    with open(os.path.join(ROOT, "profiles", "hbm_traffic_256cube.json")) as f:
        summary = json.load(f)
    t = bench.traffic_from_profiles("k_mx_pair_lds")
    assert t == summary["kernels"]["k_mx_pair_lds"]["hbm_bytes_per_launch"]
    algorithmic = 12.0 * 256 ** 3          # e_j and r32 read, e_{j+2} written
    assert algorithmic < t < 1.2 * algorithmic
    assert bench.traffic_from_profiles("no_such_kernel") is None
