"""bench.py's output contract, checked on the host: the roofline record built
from per-launch durations, the counter-measured traffic lookup and the
single-GPU guard."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from tealeaf_hip import _lib  # noqa: E402


def test_roofline_picks_dominant_kernel_and_excludes_stopped_launches():
    ur = np.concatenate([np.full(90, 0.25), np.full(10, 0.003)])  # 10 launches past a stop
    pw = np.full(100, 0.20)
    r = bench.roofline_from({_lib.TL_K_CG_UR32: ur, _lib.TL_K_CG_W32: pw}, 8192, _lib.TL_COEF_FP16)
    assert r["kernel"] == "k_cgg<4, true, 4, true, 1>"
    assert r["timed_launches"] == 90 and r["stopped_launches_excluded"] == 10
    algo = 20 * 8192 * 8192
    assert r["algorithmic_bytes_per_launch"] == algo
    assert abs(r["achieved"] - algo / 0.25e-3 / 1e9) < 0.1
    assert abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-4
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    # measured traffic is recorded for the half-coefficient kernels at 8192^2 only
    with open(bench.PMC_TRAFFIC) as f:
        rec = json.load(f)["kernels"]
    assert r["traffic"] == rec["k_cgg<4, true, 4, true, 1>"]["traffic_bytes"]
    assert bench.roofline_from({_lib.TL_K_CG_UR32: ur}, 4096)["traffic"] is None
    assert bench.roofline_from({}, 8192) is None


def test_traffic_record_covers_the_profiled_kernels():
    with open(bench.PMC_TRAFFIC) as f:
        rec = json.load(f)["kernels"]
    for kid in (_lib.TL_K_CG_W64, _lib.TL_K_CG_UR64, _lib.TL_K_CG_W32, _lib.TL_K_CG_UR32, _lib.TL_K_CGF32,
                _lib.TL_K_CGP_G32, _lib.TL_K_CGP_U32):
        name = _lib.kernel_name(kid)
        assert name in rec, name
        # counter traffic is at least the algorithmic bytes and within 15% of them
        algo = bench.KERNEL_BYTES[kid] * 8192 * 8192
        assert algo <= rec[name]["traffic_bytes"] < 1.15 * algo, (name, rec[name]["traffic_bytes"], algo)


def test_pair_summary_and_roofline_of_the_pair_form():
    """With the pair form the dominant kernel is its second sweep (26 B/cell);
    the pair record adds both sweeps: 36 B/cell per two iterations."""
    g = np.concatenate([np.full(40, 0.14), np.full(5, 0.002)])
    u = np.concatenate([np.full(40, 0.34), np.full(5, 0.003)])
    prof = {_lib.TL_K_CGP_G32: g, _lib.TL_K_CGP_U32: u}
    r = bench.roofline_from(prof, 8192, _lib.TL_COEF_FP16)
    assert r["kernel"] == "k_cgp<1, true>" and r["bytes_per_cell"] == 26 and r["timed_launches"] == 40
    assert r["traffic"] is not None
    pr = bench.pair_summary(prof, 8192)
    assert pr["first_sweep_ms"] == 0.14 and pr["second_sweep_ms"] == 0.34
    assert pr["bytes_per_cell_per_iteration"] == 18 and pr["ms_per_iteration"] == 0.24
    assert abs(pr["achieved_GBps"] - 36 * 8192 * 8192 / 0.48e-3 / 1e9) < 0.1
    assert bench.pair_summary({_lib.TL_K_CGF32: np.full(4, 0.28)}, 8192) is None


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "--gpus must be 1" in p.stderr
    assert p.stdout == ""


def test_cpu_baseline_record():
    """The CPU leg's record on a tiny grid: the 1-thread figure is the value
    (the reference is single-threaded), the all-core figure, the CPU model and
    the reference deck timed in full ride along."""
    r = bench.cpu_baseline(48, 1001, 500, 1.0, c1_n=24)
    assert r["cores"] == 1 and r["kind"] == "port" and r["unit"] == "s/timestep" and r["value"] > 0
    assert r["cpu_model"] and r["host_cores_available"] >= 1
    assert r["all_cores"]["threads"] == r["host_cores_available"]
    c1 = r["c1_24_10_steps"]
    assert c1[0]["threads"] == r["host_cores_available"] and len(c1[0]["iterations"]) == 10
    if len(c1) > 1:
        assert c1[1]["threads"] == 1 and c1[1]["temp_step10"] == c1[0]["temp_step10"]
