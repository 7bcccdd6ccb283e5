"""bench.py's contract pieces that run without a GPU: the CPU-baseline object
(all-cores oracle with the one-thread rate beside it), the traffic figure read from
the committed rocprofv3 summary, and the refusal of more than one GPU."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_cpu_baseline_object():
    from mpm_snow_sim2d_amd.params import snow_params
    from mpm_snow_sim2d_amd.scenes import snowball_disks
    P = snow_params(2, 128, 1e-4)
    ps = snowball_disks(seed=0)
    out = bench.cpu_baseline(P, ps, 1, budget_s=0.2)
    assert set(out) >= {"value", "unit", "cores", "kind", "sample", "single_thread"}
    assert out["kind"] == "port" and out["unit"] == "M particle-substeps/s"
    assert out["value"] > 0 and out["cores"] >= 1
    assert out["single_thread"]["cores"] == 1 and out["single_thread"]["value"] > 0
    # the reported value is the faster of the two legs, with that leg's thread count
    legs = (out["single_thread"], out["all_cores"])
    best = max(legs, key=lambda g: g["value"])
    assert out["value"] == best["value"] and out["cores"] == best["cores"]
    json.dumps(out)
    # the reference CPU path's own algorithm (APIC + explicit grid force) as the
    # reported legs, the MLS restatement beside them
    a = bench.cpu_baseline(P, ps, 1, budget_s=0.2, apic=True)
    assert a["value"] > 0 and a["single_thread"]["value"] > 0 and a["single_thread"]["cores"] == 1
    assert "APIC" in a["sample"] and a["glsl_mls"]["value"] > 0
    json.dumps(a)


def test_traffic_from_committed_profile():
    for superblock in (False, True):
        t = bench.traffic_from_profiles(superblock)
        assert t is not None
        # the summary's traffic is the scaled FETCH_SIZE + WRITE_SIZE of the fused
        # kernel: at least the 112 B/particle a fused substep must move (x, F_E, J_P,
        # m, V read, 60 B; x, F_E, J_P written, 52 B: the gather recomputes v and C,
        # which only a substep before a binning or the step's end stores), and not
        # absurdly more than the 200 B/particle algorithmic figure
        assert 2_000_376 * 112 <= t <= 4 * 2_000_376 * 200


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
