"""Batched scenes, graph replay and the autotuner, through the C-ABI.

A batch (mpm_batch_create) holds independent scenes with their own grids, time
steps and particle counts; every scene must evolve as it does alone.  Each scene is
compared with the oracle (FP32 and FP64) under the whole-run bound of the
single-scene parity tests (max 1e-3, median 1e-5; x in cells: max 1e-3, median
1e-4) and with a single-scene solver run of the same scene.  Graph replay must give
what eager stepping gives, checked after every replay with eager work and
synchronisations in between.
"""
import numpy as np
import pytest

from mpm_snow_sim2d_amd.params import BOUNDARY_COULOMB_MARGIN, Options, snow_params
from mpm_snow_sim2d_amd.scenes import ParticleSet, random_scene
from tests.helpers import make_case, state_errors, two_balls

pytestmark = pytest.mark.gpu

BOUND_MAX = 1e-3
BOUND_MED = {"x": 1e-4, "v": 1e-5, "C": 1e-5, "F": 1e-5, "Jp": 1e-5}


def _flush(msg):
    print(msg, flush=True)


def _batch_2d(boundary=None):
    """Six scenes: 128^2 and 256^2 grids with different time steps, one to three
    shapes, uneven counts, and one scene with no particles."""
    spec = [(128, 1e-4, 11, 1500), (256, 5e-5, 12, 4000), (128, 1e-4, 13, 800),
            (256, 5e-5, 14, 2500), (128, 1e-4, 15, 0), (128, 1e-4, 16, 3000)]
    out = []
    for cells, dt, seed, n in spec:
        P = snow_params(2, cells, dt)
        if boundary is not None:
            P.boundary = boundary
        ps = random_scene(seed, n) if n else ParticleSet.empty(2)
        out.append((P, ps))
    return out


def _batch_3d():
    out = []
    for cells, dt, seed in ((32, 5e-5, 1), (40, 5e-5, 2), (24, 1e-4, 3)):
        P = snow_params(3, cells, dt)
        out.append((P, two_balls(3, P, seed, 0.7, 3.0, 4.0)))
    return out


BATCHES = {"2d_sticky": lambda: _batch_2d(),
           "2d_coulomb": lambda: _batch_2d(BOUNDARY_COULOMB_MARGIN),
           "3d_sticky": _batch_3d}


def _assert_bound(err, tag):
    for k in ("x", "v", "C", "F", "Jp"):
        assert err[k][0] < BOUND_MAX, (tag, k, err[k])
        assert err[k][1] < BOUND_MED[k], (tag, k, err[k])


@pytest.mark.parametrize("batch", list(BATCHES))
def test_batch_matches_oracle_per_scene(orc, batch):
    from mpm_snow_sim2d_amd.solver import BatchSolver
    scenes = BATCHES[batch]()
    nsub = 20
    _flush(f"[batch oracle] {batch} counts={[ps.n for _, ps in scenes]} steps={nsub}")
    with BatchSolver(scenes, Options(sort_every=8, tile_particles=256)) as b:
        assert b.num_scenes == len(scenes)
        b.Step(nsub)
        assert b.check_finite() == 0
        st = b.get_state()
        for k, (P, ps) in enumerate(scenes):
            first, count = b.scene_range(k)
            assert count == ps.n
            if ps.n == 0:
                continue
            gpu = b.scene_state(st, k)
            o32 = orc.to_solver_state(P, orc.run(P, orc.OracleState.from_particles(ps, np.float32), nsub))
            o64 = orc.to_solver_state(P, orc.run(P, orc.OracleState.from_particles(ps, np.float64), nsub))
            e32, e64 = state_errors(gpu, o32, P.dx), state_errors(gpu, o64, P.dx)
            _flush(f"    scene {k} n={ps.n}: vs f32 x={e32['x']} F={e32['F']}")
            _assert_bound(e32, f"scene {k} vs f32")
            _assert_bound(e64, f"scene {k} vs f64")


def test_batch_matches_single_scene_runs():
    from mpm_snow_sim2d_amd.solver import BatchSolver, GPUSolver
    scenes = _batch_2d()
    nsub = 60
    _flush(f"[batch vs single] steps={nsub}")
    opts = Options(sort_every=16, tile_particles=512)
    with BatchSolver(scenes, opts) as b:
        b.Step(nsub)
        st = b.get_state()
    for k, (P, ps) in enumerate(scenes):
        if ps.n == 0:
            continue
        s = GPUSolver(P, ps, opts)
        s.Step(nsub)
        one = s.get_state()
        s.close()
        err = state_errors(_slice(st, scenes, k), one, P.dx)
        _flush(f"    scene {k}: x={err['x']} v={err['v']}")
        _assert_bound(err, f"scene {k}")


def _slice(st, scenes, k):
    a = sum(ps.n for _, ps in scenes[:k])
    c = scenes[k][1].n
    return {key: v[..., a:a + c] for key, v in st.items()}


@pytest.mark.parametrize("which", ["single", "batch"])
def test_graph_replay_matches_eager(which):
    from mpm_snow_sim2d_amd.solver import BatchSolver, GPUSolver
    opts = Options(sort_every=4, tile_particles=256)
    period = 2 * opts.sort_every
    if which == "single":
        P, ps = make_case("snow2d_quad_sticky")
        mk = lambda: GPUSolver(P, ps, opts)
        dx = P.dx
    else:
        scenes = _batch_2d()
        mk = lambda: BatchSolver(scenes, opts)
        dx = 1.0 / 256
    _flush(f"[graph] {which} period={period}")
    eager, graph = mk(), mk()
    graph.set_graph(True)
    # whole periods (replays), a remainder, then replays again from an unaligned start
    for chunk in (2 * period, period, period + 3, period, 5, 2 * period):
        eager.Step(chunk)
        graph.Step(chunk)
        a, b = eager.get_state(), graph.get_state()
        err = state_errors(b, a, dx)
        _flush(f"    after {chunk}: x={err['x']} v={err['v']}")
        _assert_bound(err, f"chunk {chunk}")
        assert graph.check_finite() == 0
    eager.close()
    graph.close()


def test_autotune_picks_candidate_and_runs():
    from mpm_snow_sim2d_amd.solver import BatchSolver, autotune
    scenes = _batch_2d()
    _flush("[autotune]")
    best, ms = autotune(scenes, trial_substeps=16, graph=True)
    _flush(f"    best={best} ms/substep={ms:.4f}")
    assert best.tile_particles in (128, 256, 512)
    assert best.sort_every in (4, 8, 16, 32, 64)
    assert best.fused == 1 and ms > 0
    with BatchSolver(scenes, best, graph=True) as b:
        b.Step(2 * best.sort_every + 1)
        assert b.check_finite() == 0


def test_batch_rejects_single_scene_calls():
    from mpm_snow_sim2d_amd._lib import MPM_E_STATE, MPMError
    from mpm_snow_sim2d_amd.solver import BatchSolver
    scenes = _batch_2d()[:2]
    with BatchSolver(scenes, Options(fused=0)) as b:
        with pytest.raises(MPMError) as ei:
            b.Step(1)
        assert ei.value.code == MPM_E_STATE
        lib = b._lib
        for fn in (lib.mpm_p2g, lib.mpm_grid_update, lib.mpm_g2p):
            assert fn(b._h) == MPM_E_STATE
