"""Benchmark of the plain 3D snow substep on one MI355X.

Workload (the configuration the headline metric is quoted on): a 0.3 m cube of
2,000,376 seeded-jittered particles (126^3 lattice) dropped from 0.5 m, 256^3 grid,
dx = 1/256, dt = 5e-5, quadratic MLS-MPM, sticky walls, Stomakhin snow parameters,
FP32 state.  A step is one substep (grid update + fused G2P/P2G, plus the periodic
binning).  Inputs are resident in HBM when the timed region starts.  The binning
period and work-item size come from the autotuner (short untimed trials on the
scene's state halfway through the timed substeps; --no-autotune uses --fused /
--sort-every / --tile); none changes results.

  python bench.py [--steps K] [--warmup W] [--no-cpu-baseline]
  python bench.py --workload large|batch|snowballs2d [...]

The other workloads print a line of the same form: `large` (16M particles in four
colliding spheres, 512^3 grid, dt 2.5e-5), `batch` (64 independent 2D scenes of
5k-80k particles on 128^2/256^2 grids in one solver, hipGraph replay, options from
the autotuner unless --no-autotune) and `snowballs2d` (two 2D snowballs, ~20k
particles, 128^2 grid, dt 1e-4, hipGraph replay).  Where the timed substeps replay
hipGraphs, the kernel durations come from a second, eager pass of the same
substeps from the same state (events are not recorded inside the graphs).

Prints one JSON line.  `roofline` prices the dominant kernel (the fused G2P+P2G
pass) at 200 algorithmic bytes per particle in 3D, 104 in 2D (FP32 read + write
of x, v, C, F_E, J_P) against the MI355X HBM peak; its duration comes from HIP
events around every fourth launch on the solver's stream over the timed region (only
that kernel is instrumented there, and a sample of its launches, since each event
pair also holds the stream for the launch's dispatch; the grid update and binning
are timed in a further pass).  `cpu_baseline` times the FP32 CPU oracle (the same algorithm) on the same
scene: the OpenMP build on every host core (the reported value) and the serial
oracle on one thread beside it, each for about 6 s of substeps.
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
# the fused kernel's HIP events in the timed region: one pair per this many launches
EVENT_EVERY = 4
BYTES_PER_PARTICLE = {3: 200.0,  # 3D: (12 + 12 + 36 + 36 + 4) B read + the same written
                      2: 104.0}  # 2D: (8 + 8 + 16 + 16 + 4) B read + the same written


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles(superblock: bool = False) -> float | None:
    """HBM bytes per launch of the fused kernel (the per-block kernel, or with
    superblock the super-block kernel) from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration)."""
    name = "g2p2g_superblock_falling_cube_traffic.txt" if superblock else "g2p2g_falling_cube_traffic.txt"
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def cpu_baseline(params, ps, substeps: int = 1, what: str | None = None, budget_s: float = 6.0,
                 apic: bool = False) -> dict:
    """The FP32 CPU oracle on the host, each leg for as many substeps as fit in about
    `budget_s` seconds after one untimed substep (at least `substeps`): the serial
    oracle on one thread (the reference's CPU path is serial), then the OpenMP build
    of the same body on every host core.  The faster leg is the reported value (on
    small scenes the threads' start-up and grid sweeps outweigh their work, so one
    thread can win), both legs are listed.  With apic the reported legs run the
    reference CPU path's own algorithm (APIC with the explicit grid force,
    src/cpu/Particles.cpp:115-131, Grid.cpp:118-184) and the GLSL path's MLS
    restatement is timed beside it; otherwise the MLS restatement (the algorithm the
    GPU runs) is reported."""
    import copy
    from mpm_snow_sim2d_amd.params import SCHEME_APIC_FORCE
    from oracle import oracle as orc
    scene = f"{what or 'the full scene'} ({ps.n} particles, grid {params.WIDTH}^{params.dim})"

    def leg(run, P):
        st = orc.OracleState.from_particles(ps, np.float32)
        run(P, st, 1)                               # first touch of the grid pages
        t0 = time.perf_counter()
        run(P, st, 2)
        per = (time.perf_counter() - t0) / 2
        k = max(substeps, min(4096, int(budget_s / max(per, 1e-6))))
        t0 = time.perf_counter()
        run(P, st, k)
        dt = time.perf_counter() - t0
        return ps.n * k / dt / 1e6, k, dt

    def legs(P):
        one, k1, dt1 = leg(orc.run, P)
        many, kn, dtn = leg(orc.run_omp, P)
        return ({"value": one, "cores": 1, "substeps": k1, "seconds": round(dt1, 2)},
                {"value": many, "cores": threads, "substeps": kn, "seconds": round(dtn, 2)})

    def report(single, all_cores, what_alg):
        best = all_cores if all_cores["value"] >= single["value"] else single
        how = f"OpenMP on {threads} threads" if best is all_cores else "1 thread (faster than all cores here)"
        return {"value": best["value"], "unit": "M particle-substeps/s", "cores": best["cores"], "kind": "port",
                "sample": f"{what_alg}, {how}, {best['substeps']} substeps of {scene} in {best['seconds']:.1f} s; "
                          f"legs: 1 thread {single['value']:.3g}, {threads} threads {all_cores['value']:.3g} "
                          f"M particle-substeps/s",
                "single_thread": single, "all_cores": all_cores}

    orc.load()
    threads = orc.omp_threads()
    s_mls, a_mls = legs(params)
    mls = report(s_mls, a_mls, "FP32 CPU oracle (GLSL-order restatement)")
    if not apic:
        return mls
    Pa = copy.deepcopy(params)
    Pa.scheme = SCHEME_APIC_FORCE
    s_ap, a_ap = legs(Pa)
    out = report(s_ap, a_ap, "FP32 CPU oracle of the reference CPU path's algorithm (APIC + explicit grid force, "
                             "src/cpu)")
    out["glsl_mls"] = {k: mls[k] for k in ("value", "cores", "single_thread", "all_cores")}
    return out


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


def learned(args) -> dict:
    """Learned-vs-plain: train the grid U-Net on trajectories of colliding 2D scenes
    (256^2 grid, ~60k particles each), then roll held-out scenes out for
    `--rollout` substeps in macro steps of 8 with the network (plain substeps where
    the residual test fails) and with the plain solver alone."""
    from mpm_snow_sim2d_amd.params import Options, snow_params
    from mpm_snow_sim2d_amd.scenes import colliding_scene
    from mpm_snow_sim2d_amd.solver import autotune
    from mpm_snow_sim2d_amd.surrogate import (ADVANCE_BALLISTIC, ADVANCE_NETWORK, ADVANCE_PLAIN, Surrogate, fit,
                                              generate, rollout, train_streaming)

    cells, dt, k, npart = 256, 5e-5, 8, 60_000
    P = snow_params(2, cells, dt)
    ntr, nte = args.train_scenes, args.test_scenes
    per = args.gen_substeps // k
    t0 = time.perf_counter()
    test = [(P, colliding_scene(5000 + i, npart)) for i in range(nte)]
    # streamed training selects its parameters on validation scenes of their own
    nva = nte if args.stream_chunks > 0 else 0
    sur = Surrogate(cells, cells, max_batch=args.batch, capacity=(ntr + nte + nva) * per, seed=0, d_ref=0.1)
    sur.set_macro(args.macro)
    sur.set_targets(args.refine)
    te = generate(sur, test, args.gen_substeps, k, first_slot=ntr * per, wall_tol=args.tol)
    te_slots = [x for ss in te for x in ss]
    va = generate(sur, [(P, colliding_scene(7000 + i, npart)) for i in range(nva)], args.gen_substeps, k,
                  first_slot=(ntr + nte) * per, wall_tol=args.tol) if nva else []
    va_slots = [x for ss in va for x in ss]

    def mean_loss(fn, slots):
        return float(np.mean([fn(slots[i:i + args.batch]) for i in range(0, len(slots), args.batch)]))

    base_loss = mean_loss(sur.baseline_loss, te_slots)
    if args.stream_chunks > 0:
        # the configured data scale, streamed: chunks of ntr new trajectories each,
        # generated into the training slots and trained on before the next chunk
        steps = args.train_steps // (args.stream_chunks * args.stream_epochs)
        st = train_streaming(sur, lambda c: [(P, colliding_scene(10_000 + c * ntr + i, npart)) for i in range(ntr)],
                             args.stream_chunks, args.gen_substeps, k, first_slot=0, steps_per_chunk=steps,
                             batch=args.batch, lr=args.lr, wall_tol=args.tol,
                             eval_fn=lambda m: mean_loss(m.loss, va_slots), epochs=args.stream_epochs,
                             patience=args.patience, log=log)
        n_train, train_steps, gen_s, train_s = st["samples"], st["steps"], st["generate_s"], st["train_s"]
        trajectories, curve = st["trajectories"], st["eval"]
    else:
        train = [(P, colliding_scene(1000 + i, npart)) for i in range(ntr)]
        t0 = time.perf_counter()
        tr = generate(sur, train, args.gen_substeps, k, wall_tol=args.tol)
        gen_s = time.perf_counter() - t0
        tr_slots = [x for ss in tr for x in ss]
        n_train, train_steps, trajectories, curve = len(tr_slots), args.train_steps, ntr, []
        log(f"[learned] {len(tr_slots)} training / {len(te_slots)} held-out samples in {gen_s:.1f} s")
        t0 = time.perf_counter()
        fit(sur, tr_slots, steps=args.train_steps, batch=args.batch, lr=args.lr,
            log_every=max(args.train_steps // 10, 1), log=log)
        train_s = time.perf_counter() - t0
    val_loss = mean_loss(sur.loss, te_slots)
    log(f"[learned] trained {train_steps} steps on {n_train} samples of {trajectories} trajectories in "
        f"{train_s:.1f} s; held-out loss {val_loss:.4f} (ballistic {base_loss:.4f})")
    # the plain solver it is compared with runs the autotuner's options (binning
    # period and work-item size) with graph replay
    opts_plain, tuned_ms = autotune([test[0]], trial_substeps=64, graph=True)
    log(f"[learned] plain solver options {opts_plain} ({tuned_ms:.4f} ms per substep in the tuning trial)")
    floor_opts = Options(sort_every=32 if opts_plain.sort_every == 16 else 16,
                         tile_particles=128 if opts_plain.tile_particles == 256 else 256)

    # a macro step runs the network on 1 image (velocity), or 2 with the
    # displacement; a training step on as many per slot
    ni = 2 if args.macro else 1
    fwd_ms = sur.time(ni, iters=20)
    step_ms = sur.time(ni * args.batch, iters=5, train=True)
    flops = sur.flops(ni)

    res = {ADVANCE_PLAIN: [], ADVANCE_NETWORK: [], ADVANCE_BALLISTIC: []}
    floor = []
    # other macro forms of the same network: only a network trained with the
    # displacement image (--macro > 0) can run the forms that use it
    variants = [int(v) for v in args.macro_variants.split(",")
                if v != "" and int(v) != args.macro and (args.macro or int(v) == 0)]
    var_res = {fl: [] for fl in variants}
    for i, (Pq, ps) in enumerate(test):
        ref = rollout(None, Pq, ps, args.rollout, k, ADVANCE_PLAIN, options=opts_plain)
        res[ADVANCE_PLAIN].append({"wall": ref["wall"], "err": 0.0, "counts": ref["counts"]})
        # the same plain rollout with other binning options: only the order of the
        # float additions differs, so this is the divergence floor of the scene
        ref2 = rollout(None, Pq, ps, args.rollout, k, ADVANCE_PLAIN, options=floor_opts)
        floor.append(float(np.linalg.norm(ref2["x"] - ref["x"], axis=1).mean()))
        for mode in (ADVANCE_NETWORK, ADVANCE_BALLISTIC):
            sur.set_macro(args.macro)
            r = rollout(sur, Pq, ps, args.rollout, k, mode, tol=args.tol, options=opts_plain)
            # a rollout that a failed residual test stopped early has no error at the
            # rollout's end: it counts as failed
            err = float(np.linalg.norm(r["x"] - ref["x"], axis=1).mean()) if not r["stopped"] else float("nan")
            if r["stopped"]:
                log(f"[learned] scene {i} mode {mode}: {r['stopped']}")
            res[mode].append({"wall": r["wall"], "err": err, "counts": r["counts"], "bad": r["nonfinite"],
                              "stopped": bool(r["stopped"])})
        # the same network with the other macro-step forms (which field moves the
        # particles, which advances F_E)
        for fl in variants:
            sur.set_macro(fl)
            r = rollout(sur, Pq, ps, args.rollout, k, ADVANCE_NETWORK, tol=args.tol, options=opts_plain)
            var_res[fl].append(float(np.linalg.norm(r["x"] - ref["x"], axis=1).mean()))
        sur.set_macro(args.macro)
        log(f"[learned] scene {i}: plain {ref['wall'] * 1e3:.1f} ms; network {res[ADVANCE_NETWORK][-1]['wall'] * 1e3:.1f} ms "
            f"err {res[ADVANCE_NETWORK][-1]['err']:.2e} m uses {res[ADVANCE_NETWORK][-1]['counts']}; ballistic err "
            f"{res[ADVANCE_BALLISTIC][-1]['err']:.2e} m")
    sur.close()

    def agg(mode, key):
        return float(np.mean([r[key] for r in res[mode]]))

    t_plain, t_net = agg(ADVANCE_PLAIN, "wall"), agg(ADVANCE_NETWORK, "wall")
    net_steps = sum(r["counts"][ADVANCE_NETWORK] for r in res[ADVANCE_NETWORK])
    all_steps = sum(sum(r["counts"].values()) for r in res[ADVANCE_NETWORK])
    err_net = agg(ADVANCE_NETWORK, "err")
    achieved = flops / (fwd_ms * 1e-3) / 1e12
    return {
        "metric": METRIC,
        "value": round(t_plain / t_net, 4),
        "unit": "x (plain wall time / learned wall time, same rollout)",
        "n_gpus": 1,
        "steps": args.rollout,
        "warmup": 0,
        "ms_per_step": round(t_net / args.rollout * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 (network, FP32 accumulation); f32 (solver)",
        "data": "synthetic",
        "config": {
            "workload": f"learned 2D surrogate: 256^2 grid, colliding scenes of ~{npart} particles (2-3 disks/boxes), "
                        f"4-level BF16 U-Net 32-64-128-128 predicting grid velocity"
                        f"{' and displacement' if args.macro else ''}, target refinement {args.refine}, "
                        f"k=8 substeps per macro step, residual tolerance {args.tol}",
            "train_trajectories": trajectories, "stream_chunks": args.stream_chunks,
            "stream_epochs": args.stream_epochs, "lr": args.lr,
            "generated_substeps_per_scene": args.gen_substeps,
            "train_samples": n_train, "train_steps": train_steps, "batch": args.batch,
            "plain_options": {"sort_every": opts_plain.sort_every, "tile_particles": opts_plain.tile_particles,
                              "hipgraph": True},
            "learned_rollout": "first macro step through mpm_surrogate_advance, the rest replayed from one "
                               "recorded hipGraph of the macro step with the residual test on the device",
            "rollouts_stopped": int(sum(r.get("stopped", False) for r in res[ADVANCE_NETWORK])),
            "generate_s": round(gen_s, 2), "train_s": round(train_s, 2),
            "heldout_loss": round(val_loss, 5), "heldout_loss_ballistic": round(base_loss, 5),
            "validation_loss_per_chunk": [round(v, 5) for v in curve],
            "test_scenes": nte, "rollout_substeps": args.rollout,
            "plain_ms": round(t_plain * 1e3, 3), "learned_ms": round(t_net * 1e3, 3),
            "mean_pos_err_m": err_net, "mean_pos_err_pct_domain": round(100 * err_net, 4),
            "within_1pct": bool(err_net <= 0.01),
            "ballistic_pos_err_m": agg(ADVANCE_BALLISTIC, "err"),
            "plain_vs_plain_pos_err_m": float(np.mean(floor)),
            "network_macro_steps_share": round(net_steps / max(all_steps, 1), 4),
            "macro_flags": args.macro, "target_refine_iters": args.refine,
            "mean_pos_err_m_by_macro_flags": {str(fl): float(np.mean(v)) for fl, v in var_res.items()},
            "forward_ms": round(fwd_ms, 4), "train_step_ms": round(step_ms, 4),
        },
        "roofline": {"bound": "mfma", "kernel": f"U-Net forward (BF16 3x3 convolutions, {ni} image(s) 256^2)",
                     "achieved": round(achieved, 3), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(achieved / 2500.0, 5), "traffic": None,
                     "flops_per_forward": flops},
        "cpu_baseline": None,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--sort-every", type=int, default=16)
    ap.add_argument("--n-side", type=int, default=126)
    ap.add_argument("--workload", choices=["cube", "large", "batch", "snowballs2d", "learned"], default="cube")
    ap.add_argument("--train-scenes", type=int, default=96)
    ap.add_argument("--test-scenes", type=int, default=4)
    ap.add_argument("--gen-substeps", type=int, default=1600)
    ap.add_argument("--train-steps", type=int, default=6000)
    ap.add_argument("--stream-chunks", type=int, default=0,
                    help="learned: train on this many chunks of --train-scenes new trajectories each, generated "
                         "one chunk at a time (--train-steps spread over them); 0: one stored set")
    ap.add_argument("--stream-epochs", type=int, default=1, help="learned: passes over the streamed chunks")
    ap.add_argument("--patience", type=int, default=0,
                    help="learned, streamed: stop after this many chunks without a lower held-out loss (0: never)")
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--rollout", type=int, default=2000)
    ap.add_argument("--tol", type=float, default=1e-3)
    ap.add_argument("--macro", type=int, default=0,
                    help="macro-step fields: 1 positions from the predicted displacement, 2 F_E from its gradient "
                         "(either: the network also predicts the displacement, a second image per sample)")
    ap.add_argument("--refine", type=int, default=0,
                    help="Jacobi steps of the least-squares fit of the target fields (0: the transfers)")
    ap.add_argument("--macro-variants", default="",
                    help="comma-separated further macro flags to roll the held-out scenes out with")
    ap.add_argument("--tile", type=int, default=512)
    ap.add_argument("--fused", type=int, default=1,
                    help="with --no-autotune: 1 per-block work items, 2 / 3 super-block work items (3D) with a "
                         "drift margin of 0 / 1 cells")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-substeps", type=int, default=1)
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

    if args.workload == "learned":
        out = learned(args)
        if not args.no_cpu_baseline:
            from mpm_snow_sim2d_amd.params import snow_params
            from mpm_snow_sim2d_amd.scenes import colliding_scene
            log("[bench] CPU baseline ...")
            out["cpu_baseline"] = cpu_baseline(snow_params(2, 256, 5e-5), colliding_scene(5000, 60_000),
                                               args.cpu_substeps, "a held-out scene, plain solver")
        print(json.dumps(out), flush=True)
        return 0

    from mpm_snow_sim2d_amd import _lib
    from mpm_snow_sim2d_amd.params import Options, snow_params
    from mpm_snow_sim2d_amd.scenes import LARGE_SEED, batch_scenes, colliding_spheres, falling_cube, snowball_disks
    from mpm_snow_sim2d_amd.solver import BatchSolver, GPUSolver, autotune

    wl = args.workload
    graph = False
    tuned = None
    if wl == "cube":
        P = snow_params(3, 256, 5e-5)
        ps = falling_cube(seed=0, n_side=args.n_side)
        desc = ("plain 3D snow: 0.3 m cube of 2,000,376 jittered particles dropped from 0.5 m, 256^3 grid, "
                "dx=1/256, dt=5e-5, quadratic MLS-MPM, sticky walls")
        grid = "256^3"
    elif wl == "large":
        P = snow_params(3, 512, 2.5e-5)
        ps = colliding_spheres(seed=LARGE_SEED)
        desc = ("plain 3D large: four colliding spheres (r 0.08-0.15, moving towards their centroid at 2-4 m/s, "
                "components within +-4 m/s; seed 2: first contact after ~157 substeps, three pairs within 1000), "
                f"{ps.n} particles, 512^3 grid, dx=1/512, dt=2.5e-5, quadratic MLS-MPM, sticky walls")
        grid = "512^3"
    elif wl == "snowballs2d":
        P = snow_params(2, 128, 1e-4)
        ps = snowball_disks(seed=0)
        graph = True   # ~20k particles: launch-bound without graph replay
        desc = (f"plain 2D snow: two snowballs r=0.08 at (0.3,0.5) and (0.7,0.55), +-3 m/s, {ps.n} particles, "
                "128^2 grid, dx=1/128, dt=1e-4, quadratic MLS-MPM, sticky walls, hipGraph replay")
        grid = "128^2"
    else:
        scenes = batch_scenes(64, seed=0)
        P = scenes[0][0]
        graph = True
        desc = (f"batched 2D: 64 independent random scenes (1-3 disks/boxes), "
                f"{sum(p.n for _, p in scenes)} particles, 5k-80k per scene, 128^2 (<20k) or 256^2 grids, "
                "hipGraph replay, quadratic MLS-MPM, sticky walls")
        grid = "128^2/256^2 per scene"
    dim = P.dim

    if wl == "batch":
        if args.no_autotune:
            opts = Options(sort_every=args.sort_every, fused=1, tile_particles=args.tile)
        else:
            opts, tuned_ms = autotune(scenes, trial_substeps=64, graph=True)
            tuned = {"tile_particles": opts.tile_particles, "sort_every": opts.sort_every,
                     "ms_per_substep": round(tuned_ms, 5)}
            log(f"[bench] autotuned {tuned}")
        s = BatchSolver(scenes, opts)
        n = s.n
    else:
        if args.no_autotune:
            opts = Options(sort_every=args.sort_every, fused=args.fused, tile_particles=args.tile)
        else:
            # pipeline, binning period and tile size from the autotuner (short trials
            # before the run, not timed; results do not depend on them).  Its trials
            # start from the state the scene reaches halfway through the timed
            # substeps (an untimed pre-run with default options), so that they see
            # the run's typical velocities: how far particles drift between binnings
            # decides the best binning period, and a body at rest has not started to.
            tune_ps = ps
            if wl in ("cube", "large"):
                with GPUSolver(P, ps, Options(sort_every=16, fused=1, tile_particles=512)) as pre:
                    pre.Step(args.warmup + args.steps // 2)
                    mid = pre.get_state()
                tune_ps = type(ps)(mid["x"], mid["v"], ps.mass, ps.vol)
                log(f"[bench] autotuning on the state after {args.warmup + args.steps // 2} substeps")
            opts, tuned_ms = autotune([(P, tune_ps)], trial_substeps=64, graph=graph)
            tuned = {"tile_particles": opts.tile_particles, "sort_every": opts.sort_every, "fused": opts.fused,
                     "ms_per_substep": round(tuned_ms, 5)}
            log(f"[bench] autotuned {tuned}")
        s = GPUSolver(P, ps, opts)
        n = ps.n
    log(f"[bench] {wl}: {n} particles ready")
    init = s.get_state()
    s.Step(args.warmup)
    s.sync()
    log(f"[bench] warmup {args.warmup} substeps done")
    s.set_state(**init)
    s.sync()

    def kernel_pass(steps, kernels=None, every=1):
        s.set_timing(True, kernels)
        s.set_timing_every(every)
        s.reset_timers()
        s.sync()
        t0 = time.perf_counter()
        s.Step(steps)
        s.sync()
        wall = time.perf_counter() - t0
        t = {k: s.kernel_time(getattr(_lib, "KERNEL_" + k)) for k in ("G2P2G", "GRID_UPDATE", "SORT", "P2G")}
        s.set_timing(False)
        return wall, t

    if graph:
        # the timed region replays graphs, which the per-kernel events do not see; a
        # separate eager pass of the same substeps, from the same state, times the kernels
        s.set_graph(True)
        s.sync()
        t0 = time.perf_counter()
        s.Step(args.steps)
        s.sync()
        wall = time.perf_counter() - t0
        s.set_graph(False)
        s.set_state(**init)
        _, kt = kernel_pass(args.steps)
    else:
        # the timed region instruments only the fused kernel, one event pair per
        # EVENT_EVERY launches (each pair brackets its launch's dispatch too, so
        # fewer of them leave the timed substeps less perturbed; the sampled
        # launches cover every phase of the binning period)
        wall, kt = kernel_pass(args.steps, kernels=[_lib.KERNEL_G2P2G], every=EVENT_EVERY)
        # durations of the other kernels (log and informative fields) from a further
        # eager pass after the timed region with every kernel instrumented
        _, kt_all = kernel_pass(min(args.steps, 256))
        for key in ("GRID_UPDATE", "SORT", "P2G"):
            kt[key] = kt_all[key]
    (ms_k, nk), (ms_gu, ngu), (ms_sort, nsort), (ms_p2g, np2g) = (kt["G2P2G"], kt["GRID_UPDATE"], kt["SORT"],
                                                                  kt["P2G"])
    bad = s.check_finite()
    s.close()
    log(f"[bench] timed {args.steps} substeps in {wall:.3f} s; g2p2g {nk} launches {ms_k:.2f} ms; "
        f"grid {ngu} launches {ms_gu:.2f} ms; sort {nsort} {ms_sort:.2f} ms; p2g {np2g} {ms_p2g:.2f} ms; bad={bad}")

    value = n * args.steps / wall / 1e6
    avg_k = ms_k / max(nk, 1) / 1e3
    achieved = n * BYTES_PER_PARTICLE[dim] / avg_k / 1e9 if nk else None
    if opts.fused >= 2:
        kname = f"k_super<2,sticky,G2P,margin {opts.fused - 2}> (fused G2P+P2G, 2x2x2-block work items)"
    else:
        kname = f"k_tiled<{dim},2,sticky,G2P{',batch' if wl == 'batch' else ''}> (fused G2P+P2G)"
    config = {"workload": desc, "particles": n, "grid": grid, "sort_every": opts.sort_every,
              "tile_particles": opts.tile_particles, "work_items": "super-block" if opts.fused >= 2 else "block",
              "hipgraph": graph, "nonfinite_particles": bad}
    if tuned:
        config["autotuned"] = tuned
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
        "config": config,
        "roofline": {"bound": "hbm", "kernel": kname,
                     "achieved": round(achieved, 2) if achieved else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5) if achieved else None,
                     "traffic": traffic_from_profiles(opts.fused >= 2) if wl == "cube" else None,
                     "avg_launch_us": round(avg_k * 1e6, 3),
                     "timed_launches": nk,
                     "event_every": 1 if graph else EVENT_EVERY,
                     "grid_update_avg_us": round(ms_gu / max(ngu, 1) * 1e3, 3),
                     "sort_avg_us": round(ms_sort / max(nsort, 1) * 1e3, 3)},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline ...")
        if wl == "batch":
            big = max(scenes, key=lambda t: t[1].n)
            out["cpu_baseline"] = cpu_baseline(big[0], big[1], args.cpu_substeps, "the largest scene of the batch")
        else:
            # the 2D two-snowball scene is the reference CPU path's own case: its
            # algorithm (APIC + explicit force) is the reported baseline there
            out["cpu_baseline"] = cpu_baseline(P, ps, args.cpu_substeps, apic=(wl == "snowballs2d"))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
