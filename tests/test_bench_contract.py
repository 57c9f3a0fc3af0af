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


def test_recorded_traffic_parses_and_host_threads_follow_the_cpu_share(monkeypatch):
    import bench
    t, src = bench.measured_traffic("p3m_disk_1M_128")
    assert t is not None and 1e8 < t < 1e11 and src.startswith("profiles/")
    assert bench.measured_traffic("no_such_workload") == (None, None)
    monkeypatch.setenv("OMP_NUM_THREADS", "3")
    n, counts = bench.host_threads()
    assert n == min(3, counts["affinity"]) and counts["nproc"] >= counts["affinity"] >= 1
    monkeypatch.delenv("OMP_NUM_THREADS")
    assert bench.host_threads()[0] == counts["affinity"]


def test_workload_configs_name_the_error_target_workload():
    import bench
    from particlesimulation_amd import ics
    c = bench.workload_config(ics.C2_P3M_DISK_ERR, 42)
    assert c["interpolation"] == "TSC_CENTRED" and c["finite_diff"] == "FOUR_POINT"
    assert c["a_cells"] == c["cutoff_cells"] == 3.5 and c["n"] == 1048576 and c["grid"] == [128, 128, 128]
    c = bench.workload_config(ics.C2_P3M_DISK_ERR_REF, 42)
    assert c["interpolation"] == "TSC" and c["finite_diff"] == "TWO_POINT" and c["a_cells"] == 5.0
    json.dumps(bench.workload_config(ics.C2_P3M_DISK, 42))


def test_traffic_file_matches_its_own_counters():
    with open(os.path.join(ROOT, "profiles", "k_sr_p3m_disk_1M_128_traffic.json")) as f:
        d = json.load(f)
    bytes_ = d["fetch_kib_per_dispatch"] * 1024 * d["fetch_scale"] + d["write_kib_per_dispatch"] * 1024 * d["write_scale"]
    assert abs(bytes_ - d["traffic_bytes_per_launch"]) <= 1e-6 * bytes_


def test_cpu_baseline_samples_every_cell_and_scales_by_pairs():
    """The short-range sample takes every 4th particle of each cell's list as
    the home particle: about a quarter of the pass's ordered pairs, scaled
    back by the whole pass's pair count.  A stride of 1 is the whole pass."""
    import dataclasses
    import numpy as np
    import bench
    from oracle import oracle as O
    from particlesimulation_amd import ics
    w = dataclasses.replace(ics.C2_P3M_DISK, name="disk_small", n=20000, grid=(32, 32, 32), box=8.0)
    pos, _, m = w.code_units(seed=3)
    box = np.array([w.box] * 3, np.float32)
    args = (pos, m, box, w.rc_cells * w.H, w.H, w.rc_cells, w.a_cells, w.softening, w.cloud_shape)
    a_full, p_full = O.p3m_short_range(*args, use_table=2)
    a_one, p_one = O.p3m_short_range(*args, use_table=2, sample_stride=1)
    assert p_one == p_full and np.array_equal(a_one, a_full)
    _, p_s = O.p3m_short_range(*args, use_table=2, sample_stride=4)
    assert 0.22 < p_s / p_full < 0.28
    # a stride beyond every cell's count keeps one home particle per cell
    _, p_first = O.p3m_short_range(*args, use_table=2, sample_stride=1 << 30)
    assert 0 < p_first < p_s
    c = bench.cpu_baseline(w, pos, m, sr_pairs=p_full)
    assert c["sr_pairs"] == p_full and c["sr_sample_stride"] == 4 and c["sr_sample_pairs"] == p_s
    assert c["sr_s"] >= c["sr_fixed_s"] > 0 and c["s_per_step"] == c["pm_s"] + c["sr_s"]
    assert c["kind"] == "port" and "every 4th particle" in c["sample"]
    json.dumps(c)


class _RecordingContext:
    """Stands in for particlesimulation_amd.PM in bench.measure: records the
    calls and answers with fixed stage times (no GPU)."""

    def __init__(self, escape_on_run=None):
        self.calls, self.level, self.runs, self.escape_on_run = [], 0, 0, escape_on_run

    def set_particles(self, *a):
        self.calls.append(("set_particles",))

    def run(self, n, initial=False, p3m=False, sync=False):
        self.runs += 1
        self.calls.append(("run", n, bool(initial), self.level))
        return n

    def escaped(self):
        return self.escape_on_run is not None and self.runs >= self.escape_on_run

    def profile(self, on=True, short_range_only=False):
        self.level = 0 if not on else (2 if short_range_only else 1)
        self.calls.append(("profile", self.level))

    def profile_reset(self):
        pass

    def profile_read(self):
        if self.level == 2:
            return {"sr_pairs": (50.0, 5)}
        return {"sr_pairs": (50.0, 5), "deposit": (1.0, 5), "interp": (0.5, 5), "sr_sort": (2.0, 5)}

    def synchronize(self):
        pass

    def p3m_count_pairs(self, tested=False):
        return (1000, 1200) if tested else 1000

    def get_tuning(self):
        return (0, 0, 32, 6, 0, 0)

    def p3m_cell_geometry(self):
        return (48, 18, 1, 1000)


def test_measure_times_the_value_with_short_range_events_only_and_restarts_each_region():
    import bench
    from particlesimulation_amd import ics
    w = ics.C2_P3M_DISK
    pm = _RecordingContext()
    res = bench.measure(pm, w, 5, 2, None, None, None)
    runs = [c for c in pm.calls if c[0] == "run"]
    # three regions, each from the initial particles and the warm-up: the
    # value's run with the short-range events only, every stage bracketed,
    # then the graph's capture (4 untimed iterations) and a run with no events
    assert [c for c in pm.calls if c[0] == "set_particles"] == [("set_particles",)] * 3
    assert runs == [("run", 2, True, 0), ("run", 5, False, 2), ("run", 2, True, 0), ("run", 5, False, 1),
                    ("run", 2, True, 0), ("run", 4, False, 0), ("run", 5, False, 0)]
    assert res["roofline"]["avg_launch_ms"] == 10.0 and res["roofline"]["ordered_pairs"] == 1000
    assert set(res["stage_ms_per_step"]) == {"sr_pairs", "deposit", "interp", "sr_sort"}
    assert res["iterations"] == 5 and res["escaped"] is False
    assert res["with_stage_events"] and res["without_events"]


def test_measure_exits_without_a_line_when_the_timed_run_escaped():
    import pytest
    import bench
    from particlesimulation_amd import ics
    pm = _RecordingContext(escape_on_run=2)
    with pytest.raises(SystemExit) as e:
        bench.measure(pm, ics.C2_P3M_DISK, 5, 2, None, None, None)
    assert e.value.code == 3


def test_measure_times_a_pm_workload_on_the_graph_without_events():
    import bench
    from particlesimulation_amd import ics
    pm = _RecordingContext()
    res = bench.measure(pm, ics.C3_PM_CLUMPS, 5, 2, None, None, None)
    runs = [c for c in pm.calls if c[0] == "run"]
    assert runs == [("run", 2, True, 0), ("run", 4, False, 0), ("run", 5, False, 0), ("run", 2, True, 0),
                    ("run", 5, False, 1)]
    assert res["roofline"]["bound"] == "hbm" and res["roofline"]["kernel"] in ("deposit", "interp")
    assert res["without_events"] is None
