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
    for name in ("main", "compare_methods", "animated_kernel", "kd_build_only", "kernel_name", "profiled_traffic"):
        assert callable(getattr(b, name)), name
    assert b.kernel_name({"variant": 3, "pipeline": 2, "leaf_batch": 2}) == "k_trace<3,false,true,2>"
    assert b.kernel_name({"variant": 5, "pipeline": 0}) == "k_trace<5,false,true,1>"
    assert b.kernel_name({"variant": 5, "pipeline": 1}) == "queued pipeline graph"


def test_traffic_profile_lookup():
    b = _bench()
    lines = [l.split() for f in (ROOT / "profiles").glob("*traffic*.txt") for l in f.read_text().splitlines()
             if l.startswith("traffic ")]
    assert lines, "no traffic line in profiles/"
    kernel, opts, res, nbytes = lines[0][1], lines[0][2], lines[0][3], int(lines[0][4])
    o = {k: int(v) for k, v in (kv.split("=") for kv in opts.split(","))}
    W, H = (int(x) for x in res.split("x"))
    got, src = b.profiled_traffic(kernel, o, W, H)
    assert got == nbytes and "profiles/" in src
    assert b.profiled_traffic("k_trace<9,false,true,1>", o, W, H)[0] is None
