"""bench.py contract pieces that run without a GPU: the multi-GPU refusal,
the helpers its modes use, and the lookup of the committed traffic profile."""
from __future__ import annotations

import importlib.util
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_modes_and_helpers_exist():
    b = _bench()
    for name in ("main", "compare_methods", "run_workload", "roofline", "kernel_only", "kd_build_only",
                 "kernel_name", "profiled_counters", "host_cpus", "classify_vs_reference"):
        assert callable(getattr(b, name)), name
    assert b.kernel_name({"variant": 3, "pipeline": 2, "leaf_batch": 2}) == "k_trace<3,false,true,2>"
    assert b.kernel_name({"variant": 5, "pipeline": 0}) == "k_trace<5,false,true,1>"
    assert b.kernel_name({"variant": 5, "pipeline": 1}) == "queued pipeline graph"
    assert b.kernel_name({"variant": 5, "pipeline": 3, "leaf_batch": 2}) == "k_render_refill<5,false,2,false>"
    assert b.kernel_name({"variant": 5, "pipeline": 3, "leaf_batch": 2, "coop_lanes": 64}) == \
        "k_render_refill<5,false,1,true>"
    threads, aff, model = b.host_cpus()
    assert 1 <= threads <= aff and isinstance(model, str)


def test_counter_profile_lookup(tmp_path):
    """Only a line of the same kernel, options and workload counts."""
    b = _bench()
    opts = {"variant": 3, "block_size": 64, "short_stack_depth": 6, "lds_top_levels": 6, "persistent": 0,
            "pipeline": 2, "leaf_batch": 2, "coop_lanes": 0}
    (tmp_path / "profiles").mkdir()
    (tmp_path / "profiles" / "x.txt").write_text(
        f"counters k_trace<3,false,true,2> {b.opt_string(opts)} animated1M_poses10-19_1024x1024 "
        "fetch_bytes=100 write_bytes=50 valu_wave_instr=1000\n")
    b.ROOT = tmp_path
    got, src = b.profiled_counters("k_trace<3,false,true,2>", opts, "animated1M_poses10-19_1024x1024")
    assert got == {"fetch_bytes": 100.0, "write_bytes": 50.0, "valu_wave_instr": 1000.0} and src.endswith("x.txt")
    assert b.profiled_counters("k_trace<3,false,true,2>", dict(opts, block_size=128),
                               "animated1M_poses10-19_1024x1024")[0] is None
    assert b.profiled_counters("k_trace<3,false,true,2>", opts, "terrain163k_poses10-19_1024x1024")[0] is None


def test_roofline_picks_the_most_loaded_resource(tmp_path):
    b = _bench()
    opts = {"variant": 5, "block_size": 64, "short_stack_depth": 6, "lds_top_levels": 0, "persistent": 1,
            "pipeline": 3, "leaf_batch": 2, "coop_lanes": 16}
    (tmp_path / "profiles").mkdir()
    # 1 ms launch: 307.2 G wave instructions/s (half the VALU peak), 1 TB/s HBM (1/8 of peak)
    (tmp_path / "profiles" / "y.txt").write_text(
        f"counters k_render_refill<5,false,2,false> {b.opt_string(opts)} w_1024x1024 "
        "fetch_bytes=600000000 write_bytes=400000000 valu_wave_instr=307200000\n")
    b.ROOT = tmp_path
    wl = {"options_batched": opts, "timing": {"batched": (1.01, 1.0)}, "workload_key": "w_1024x1024",
          "algorithmic": {"bytes_per_launch": 20_000_000_000}}
    r = b.roofline(wl, 1024, 1024)
    assert r["bound"] == "valu_issue" and abs(r["frac"] - 0.5) < 1e-6
    assert abs(r["hbm"]["frac"] - 0.125) < 1e-6 and r["traffic"] == 1_000_000_000
    assert r["cache_served_algorithmic_GBs"] == 20000.0
    wl["workload_key"] = "other"
    assert b.roofline(wl, 1024, 1024)["frac"] is None
