"""Benchmark of the plain 3D snow substep on one MI355X.

Workload (the configuration the headline metric is quoted on): a 0.3 m cube of
2,000,376 seeded-jittered particles (126^3 lattice) dropped from 0.5 m, 256^3 grid,
dx = 1/256, dt = 5e-5, quadratic MLS-MPM, sticky walls, Stomakhin snow parameters,
FP32 state.  A step is one substep (grid update + fused G2P/P2G, plus the periodic
binning).  Inputs are resident in HBM when the timed region starts.

  python bench.py [--steps K] [--warmup W] [--no-cpu-baseline]

Prints one JSON line.  `roofline` prices the dominant kernel (the fused G2P+P2G
pass) at 200 algorithmic bytes per particle (FP32 read + write of x, v, C, F_E,
J_P) against the MI355X HBM peak; its duration comes from HIP events around each
launch on the solver's stream over the timed region.  `cpu_baseline` times the FP32
CPU oracle (the same algorithm, one host thread) on 2 substeps of the same scene.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "M particle-substeps/s, 3D 2M pts; learned-vs-plain speedup at ≤1% pos err"
HBM_PEAK_GBS = 8000.0
BYTES_PER_PARTICLE = 200.0     # 3D: (12 + 12 + 36 + 36 + 4) B read + the same written


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles() -> float | None:
    """HBM bytes per launch of the fused kernel from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration)."""
    path = os.path.join(ROOT, "profiles", "g2p2g_falling_cube_traffic.txt")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def cpu_baseline(params, ps, substeps: int = 2) -> dict:
    from oracle import oracle as orc
    st = orc.OracleState.from_particles(ps, np.float32)
    orc.load()
    t0 = time.perf_counter()
    orc.run(params, st, substeps)
    dt = time.perf_counter() - t0
    return {"value": ps.n * substeps / dt / 1e6, "unit": "M particle-substeps/s", "cores": 1,
            "kind": "port",
            "sample": f"FP32 CPU oracle (GLSL-order restatement), 1 thread, {substeps} substeps of the "
                      f"full {ps.n}-particle cube on the 256^3 grid ({dt:.1f} s)"}


def copy_calibration(gib: float = 2.0) -> None:
    """A plain device-to-device copy of `gib` GiB (larger than the 256 MB last-level
    cache), for calibrating rocprofv3's FETCH_SIZE / WRITE_SIZE against bytes moved:
    run this under `rocprofv3 --pmc FETCH_SIZE` (then WRITE_SIZE) and compare the
    __amd_rocclr_copyBuffer dispatch with the size printed here."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = int(gib * (1 << 30))
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0     # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes_each_way": nbytes, "copies": 3}), flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--sort-every", type=int, default=16)
    ap.add_argument("--n-side", type=int, default=126)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-substeps", type=int, default=2)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 2 GiB device copy (rocprofv3 counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (one process on one "
              "MI355X; the multi-GPU domain decomposition is design-only)", file=sys.stderr)
        return 2

    from mpm_snow_sim2d_amd import _lib
    from mpm_snow_sim2d_amd.params import Options, snow_params
    from mpm_snow_sim2d_amd.scenes import falling_cube
    from mpm_snow_sim2d_amd.solver import GPUSolver

    P = snow_params(3, 256, 5e-5)
    ps = falling_cube(seed=0, n_side=args.n_side)
    log(f"[bench] scene ready: {ps.n} particles")
    s = GPUSolver(P, ps, Options(sort_every=args.sort_every, fused=1, tile_particles=512))
    init = s.get_state()
    s.Step(args.warmup)
    s.sync()
    log(f"[bench] warmup {args.warmup} substeps done")
    s.set_state(**init)
    s.sync()

    s.set_timing(True)
    s.reset_timers()
    s.sync()
    t0 = time.perf_counter()
    s.Step(args.steps)
    s.sync()
    wall = time.perf_counter() - t0
    ms_k, nk = s.kernel_time(_lib.KERNEL_G2P2G)
    ms_gu, ngu = s.kernel_time(_lib.KERNEL_GRID_UPDATE)
    ms_sort, nsort = s.kernel_time(_lib.KERNEL_SORT)
    ms_p2g, np2g = s.kernel_time(_lib.KERNEL_P2G)
    bad = s.check_finite()
    s.close()
    log(f"[bench] timed {args.steps} substeps in {wall:.3f} s; g2p2g {nk} launches {ms_k:.2f} ms; "
        f"grid {ngu} launches {ms_gu:.2f} ms; sort {nsort} {ms_sort:.2f} ms; p2g {np2g} {ms_p2g:.2f} ms; bad={bad}")

    value = ps.n * args.steps / wall / 1e6
    avg_k = ms_k / max(nk, 1) / 1e3
    achieved = ps.n * BYTES_PER_PARTICLE / avg_k / 1e9 if nk else None
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "M particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain 3D snow: 0.3 m cube of 2,000,376 jittered particles dropped from "
                               "0.5 m, 256^3 grid, dx=1/256, dt=5e-5, quadratic MLS-MPM, sticky walls",
                   "particles": ps.n, "grid": "256^3", "sort_every": args.sort_every,
                   "nonfinite_particles": bad},
        "roofline": {"bound": "hbm", "kernel": "k_tiled<3,2,sticky,G2P> (fused G2P+P2G)",
                     "achieved": round(achieved, 2) if achieved else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5) if achieved else None,
                     "traffic": traffic_from_profiles(),
                     "avg_launch_us": round(avg_k * 1e6, 3),
                     "grid_update_avg_us": round(ms_gu / max(ngu, 1) * 1e3, 3),
                     "sort_avg_us": round(ms_sort / max(nsort, 1) * 1e3, 3)},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline ...")
        out["cpu_baseline"] = cpu_baseline(P, ps, args.cpu_substeps)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
