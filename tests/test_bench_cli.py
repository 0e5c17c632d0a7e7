"""bench.py without a GPU: the one-GPU contract (more GPUs exit with a message), the launch-variant
kernel names, and the HBM traffic read back from the committed rocprofv3 summary."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert msg["n_gpus"] == 2 and "one GPU" in msg["error"]


def test_kernel_names_follow_the_launch_variants():
    assert bench.kernel_name("hex", "double", 0) == "k_hex_forces<double, double, 256, 2, false>"
    assert bench.kernel_name("q4", "float", 3) == "k_q4_forces<float, float, 256, 3, false>"


def test_traffic_of_the_chosen_hex_variant_is_in_the_committed_profile():
    t = bench.measured_traffic(bench.kernel_name("hex", "double", 0))
    assert t is not None and 1.5e9 < t < 4.0e9
    assert bench.measured_traffic("k_not_a_kernel<double>") is None
