"""bench.py's host logic, without a GPU: the committed --pmc summaries parse
into the figures the bench line reports, the limiter is computed from them
the way its docstring states, and --gpus above 1 is refused (one GPU is
measured)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


@pytest.mark.parametrize("path", [bench.TRAFFIC_FILE, bench.C2_PMC_FILE])
def test_committed_profiles_hold_the_limiter_figures(path):
    p = bench.read_profile(path)
    for k in ("traffic_bytes_per_launch", "l2_hit_rate", "valu_wave_instructions_per_launch", "valu_lanes_active",
              "valu_peak_wave_instr_per_s"):
        assert k in p, (path, k)
    assert 0.0 < p["l2_hit_rate"] < 1.0 and 0.0 < p["valu_lanes_active"] <= 1.0
    assert p["traffic_bytes_per_launch"] > 1e8


def test_limiter_follows_its_definition(tmp_path):
    f = tmp_path / "pmc.txt"
    f.write_text("header: text\n"
                 "traffic_bytes_per_launch: 4000000000\n"
                 "l2_hit_rate: 0.8\n"
                 "valu_wave_instructions_per_launch: 1000000000\n"
                 "valu_lanes_active: 0.5\n"
                 "valu_peak_wave_instr_per_s: 1e12\n")
    lim = bench.limiter(str(f), 2e-3)
    assert lim["kind"] == "VALU issue"
    assert lim["valu_issue_frac"] == pytest.approx(0.5)        # 1e9 / 2 ms / 1e12
    assert lim["hbm_measured_gbs"] == pytest.approx(2000.0)    # 4e9 B / 2 ms
    assert lim["hbm_measured_frac"] == pytest.approx(0.25)     # of 8 TB/s
    assert lim["l2_hit_rate"] == 0.8 and lim["valu_lanes_active"] == 0.5
    assert bench.limiter(str(tmp_path / "missing.txt"), 2e-3) is None


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
