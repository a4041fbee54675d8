"""bench.py's contract pieces that need no GPU: the one-GPU refusal, the
roofline objects' arithmetic and the traffic read from the committed counter
summary."""
import importlib.util
import os
import subprocess
import sys

import pytest

import cases

ROOT = cases.ROOT


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_rp_roofline_is_flops_over_launch_time():
    b = _bench()
    # 10 launches of 1e8 sub-steps on 2^22 bubbles, 4 ms each
    sm = {"substeps": 1e9, "launches": 10, "rp_launch_ms": 4.0}
    r = b.rp_roofline(sm, 1 << 22, 1.3e9)
    flops = (b.RP_FLOPS_PER_SUBSTEP * 1e8 + b.BUBBLE_FLOPS_PER_CALL * (1 << 22))
    assert r["bound"] == "fp64-valu" and r["unit"] == "TFLOP/s"
    assert r["flops_per_launch"] == pytest.approx(flops)
    assert r["achieved"] == pytest.approx(flops / 4e-3 / 1e12)
    assert r["frac"] == pytest.approx(r["achieved"] / 78.6)
    assert r["traffic"] == 1.3e9


def test_step_and_grid_rooflines_use_the_byte_models():
    b = _bench()
    n = 4096
    st = [{"n_pic": 5, "rp_mean_substeps": 26.0, "rp_max_substeps": 3717,
           "rp_lean_launches": 5, "rp_reused": 3 << 20}] * 4
    ph = {"velocity": 0.4, "motion": 0.6, "pressure": 0.8, "bubble_rp": 90.0, "other": 1.2}
    sm = b.summarize(st, 0.1, ph, n, 1 << 22, "mixed")
    want = n * n * (116 + 56 * 5) + (1 << 22) * (96 + 288 * 5)  # per step
    assert sm["step_bytes"] == want
    assert sm["device_ms_per_step"] == pytest.approx(93.0 / 4)
    assert sm["rp_kept_fraction"] == pytest.approx(3.0 / 20.0)  # 3 * 2^20 of 5 * 2^22 calls
    r = b.step_roofline(sm, "mixed")
    assert r["bound"] == "hbm" and r["unit"] == "GB/s"
    assert r["achieved"] == pytest.approx(want / (93.0 / 4 * 1e-3) / 1e9)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0)
    g = b.grid_roofline(ph, n, 4, mixed=True)
    assert g["ms_per_step"] == pytest.approx(0.3)
    assert g["achieved"] == pytest.approx(52.0 * n * n / 0.3e-3 / 1e9)


def test_traffic_comes_from_the_committed_counter_summary():
    b = _bench()
    v = b.measured_traffic("k_bubble_persistent_4096_counters.md")
    text = open(os.path.join(ROOT, "profiles", "k_bubble_persistent_4096_counters.md")).read()
    line = [x for x in text.splitlines() if x.startswith("traffic_bytes_per_launch:")][0]
    assert v == float(line.split(":")[1]) and v > 1e9
    assert b.measured_traffic("no_such_file.md") is None
