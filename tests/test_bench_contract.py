"""bench.py's command-line contract and the committed measurements it reports
(CPU only: nothing here touches a GPU)."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_recorded_traffic_and_learned_comparison_parse():
    import bench
    t, src = bench.measured_traffic("p3m_disk_1M_128")
    assert t is not None and 1e8 < t < 1e11 and src.startswith("profiles/")
    assert bench.measured_traffic("no_such_workload") == (None, None)
    d = bench.measured_learned()
    assert d is not None and d["workload"] == "pm_clumps_4M_256"
    for k in ("plain_pm", "learned_pm", "plain_p3m"):
        assert d[k]["e_rms"] > 0 and d[k]["updates_per_s"] > 0
    json.dumps(d)  # serialisable into the bench line


def test_traffic_file_matches_its_own_counters():
    with open(os.path.join(ROOT, "profiles", "k_sr_p3m_disk_1M_128_traffic.json")) as f:
        d = json.load(f)
    bytes_ = d["fetch_kib_per_dispatch"] * 1024 * d["fetch_scale"] + d["write_kib_per_dispatch"] * 1024 * d["write_scale"]
    assert abs(bytes_ - d["traffic_bytes_per_launch"]) <= 1e-6 * bytes_
