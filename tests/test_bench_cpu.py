"""bench.py's measurement arithmetic on the CPU: the roofline entry, the
per-kernel table, the merged profiles, the counter traffic table it reads,
and the one-GPU guard.  No GPU call is made."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402  (standard library only at import time)

K26 = "kronecker s26 ef16 seed 42, degree-ordered layout"


def test_roofline_takes_the_longest_kernel_credited_with_bytes():
    prof = {"gathers": {"launches": 2, "ms": 4.0, "alg_bytes": 16e9},
            "uncredited": {"launches": 10, "ms": 10.0, "alg_bytes": 0.0},
            "short": {"launches": 1, "ms": 1.0, "alg_bytes": 1e9}}
    r = bench.roofline(prof, {"gathers": {"bytes_per_launch": 12345}}, copy_gbs=5000.0)
    assert r["kernel"] == "gathers"
    assert r["achieved"] == pytest.approx(4000.0)  # 16 GB in 4 ms
    assert r["frac"] == pytest.approx(4000.0 / bench.HBM_PEAK_GBS)
    assert r["peak"] == 8000.0 and r["unit"] == "GB/s" and r["bound"] == "hbm"
    assert r["alg_bytes_per_launch"] == 8_000_000_000
    assert r["avg_launch_us"] == pytest.approx(2000.0)
    assert r["traffic"] == 12345
    assert r["frac_of_copy"] == pytest.approx(0.8)
    assert bench.roofline(prof, {}, 5000.0)["traffic"] is None
    assert bench.roofline({"x": {"launches": 1, "ms": 1.0, "alg_bytes": 0.0}}, {}, 5000.0) is None


def test_profiles_merge_and_table_order():
    a = {"k1": {"launches": 1, "ms": 2.0, "alg_bytes": 4e9}}
    b = {"k1": {"launches": 3, "ms": 1.0, "alg_bytes": 2e9}, "k2": {"launches": 1, "ms": 5.0, "alg_bytes": 0.0}}
    m = bench.merge_reports([a, b])
    assert m["k1"] == {"launches": 4, "ms": 3.0, "alg_bytes": 6e9}
    t = bench.kernel_table(m)
    assert list(t) == ["k2", "k1"]  # by total device time
    assert t["k1"]["avg_us"] == pytest.approx(750.0) and t["k1"]["alg_gbs"] == pytest.approx(2000.0)
    assert t["k2"]["alg_gbs"] == 0.0


def test_traffic_table_covers_every_roofline_kernel():
    """Each Kronecker 26 line's roofline kernel has a counter figure, and each
    figure is FETCH_SIZE x fetch factor + WRITE_SIZE x write factor."""
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
        tj = json.load(f)
    cal = tj["calibration"]
    assert cal["fetch_factor"] == pytest.approx(2.0, rel=1e-3)
    assert cal["write_factor"] == pytest.approx(1.0, rel=1e-3)
    t = bench.load_traffic(K26)
    for k in ("msbfs_bottom_up", "bfs_bottom_up", "sssp_relax", "sssp_batch_relax", "pr_blocks_sum",
              "pr_blocks_write_fp64", "cc_link_round", "sv_hook", "bc_delta_hub", "msbc_delta_hub_wide", "tc_count"):
        e = t[k]
        assert e["bytes_per_launch"] > 0 and e["dispatches"] > 0, k
        want = (e["fetch_kib"] * cal["fetch_factor"] + e["write_kib"] * cal["write_factor"]) * 1024
        assert e["bytes_per_launch"] == pytest.approx(want, rel=1e-3), k


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout == ""
