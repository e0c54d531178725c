"""Host-side parts of bench.py (no GPU): the committed traffic summary it
reads for roofline.traffic, the one-GPU guard, and the per-workload table
of its JSON line."""
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_recorded_traffic_reads_the_committed_summary():
    import bench
    path = os.path.join(ROOT, "profiles", "stage2d_turb1024_fp64_traffic.md")
    text = open(path).read()
    for kernel in ("sweep_x", "sweep_y"):
        val, src = bench.recorded_traffic("2d_decaying_mhd_1024x1024_hlld_wcns4_ct_rk3", kernel)
        assert val == float(re.search(r"%s=(\d+)" % kernel, text).group(1))
        assert src["file"] == "profiles/stage2d_turb1024_fp64_traffic.md"
        assert re.fullmatch(r"[0-9a-f]{7,40}", src["code_commit"])
    # the algorithmic bytes of the y sweep (186.7 MB) against its measured traffic
    val, _ = bench.recorded_traffic("2d_decaying_mhd_1024x1024_hlld_wcns4_ct_rk3", "sweep_y")
    assert 1.0 <= val / 186.7e6 < 1.2
    assert bench.recorded_traffic("some_other_workload", "sweep_y") == (None, None)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_config_table_pairs_each_workload_with_its_baseline():
    import bench
    line = {"value": 1.6e9, "cpu_baseline": {"value": 5.0e6},
            "brio_wu": {"value": 2.4e7, "cpu_baseline": {"value": 1.0e5}},
            "mixed": {"fp64_value": 1.4e9, "mixed_value": 2.4e9, "l1_rho": 3.7e-8, "cpu_baseline": {"value": 4.0e6}},
            "batch": {"batched_value": 1.5e9, "cpu_baseline": {"value": 5.0e6}},
            "inverse": {"ratio": 6.1}}
    t = bench.config_table(line)
    assert sorted(t) == ["1_brio_wu_1d_1024", "2_turb2d_1024", "3_khi_2048_mlau", "4_turb3d_512", "5_batch_64"]
    assert t["2_turb2d_1024"]["gpu_over_cpu"] == 1.6e9 / 5.0e6
    assert t["3_khi_2048_mlau"]["gpu_mixed"] == 2.4e9
    assert t["4_turb3d_512"]["gpu_fp64"] is None and "gpu_over_cpu" not in t["4_turb3d_512"]
    assert t["5_batch_64"]["adjoint_gradient_over_forward"] == 6.1
