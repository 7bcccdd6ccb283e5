"""mpm_surrogate_rollout (GPU, through the C-ABI): macro steps replayed from one
recorded hipGraph with the residual test on the device against the same number of
mpm_surrogate_advance calls (same kernels on the same buffers), across repeated
calls with host work and synchronisation between them.  The capture adds its
workgroups' boxes into the image with float atomics, so two step-by-step runs
already differ at round-off (and the network's BF16 inputs can round either way
where they do): the replayed rollout may differ from one of them by at most 4x what
the two differ by, plus a floor of ~80 FP32 ulps at 1 (the largest difference over
~3000 particles and 12 macro steps is a single particle's, so the run-to-run spread
alone can exceed the factor).  A failing residual test stops
the rollout at that macro step with the particles as they were before it."""
import numpy as np
import pytest

from mpm_snow_sim2d_amd.params import Options, snow_params
from mpm_snow_sim2d_amd.scenes import colliding_scene, jitter_lattice, _set

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("mode_name", ["network", "ballistic"])
def test_graph_rollout_matches_stepwise_advance(mode_name):
    from mpm_snow_sim2d_amd.solver import GPUSolver
    from mpm_snow_sim2d_amd.surrogate import ADVANCE_BALLISTIC, ADVANCE_NETWORK, Surrogate
    mode = ADVANCE_NETWORK if mode_name == "network" else ADVANCE_BALLISTIC
    P = snow_params(2, 64, 2e-4)
    ps = colliding_scene(5, 3000)
    k = 8
    _flush(f"[rollout] graph vs stepwise, {mode_name}, n={ps.n}")
    from tests.helpers import state_errors
    floor = {"x": 1e-4, "v": 1e-5, "C": 1e-5, "F": 1e-5, "Jp": 1e-5}

    def check(sa, sb, sc):
        e_ab = state_errors(sa, sb, float(P.dx))
        e_bc = state_errors(sb, sc, float(P.dx))
        _flush(f"    graph vs stepwise {e_ab}\n    stepwise vs stepwise {e_bc}")
        for key in floor:
            assert e_ab[key][0] <= 4.0 * e_bc[key][0] + floor[key], (key, e_ab[key], e_bc[key])

    with Surrogate(64, 64, max_batch=1, capacity=0, seed=3) as sur:
        a, b, c = (GPUSolver(P, ps, Options(sort_every=4, tile_particles=256)) for _ in range(3))
        for s in (a, b, c):
            s.Step(64)                      # the bodies deform: C and F_E non-trivial
        nn, np_, res = sur.rollout_steps(a, k, 12, mode, 1.0)
        assert (nn, np_) == (12, 0) and 0.0 <= res <= 1.0
        for s in (b, c):
            for _ in range(12):
                used, _ = sur.advance(s, k, mode, 1.0)
                assert used == mode
        check(a.get_state(), b.get_state(), c.get_state())
        # a second call replays the recorded graph (host reads and syncs in between)
        nn, np_, _ = sur.rollout_steps(a, k, 5, mode, 1.0)
        assert (nn, np_) == (5, 0)
        for s in (b, c):
            for _ in range(5):
                sur.advance(s, k, mode, 1.0)
        check(a.get_state(), b.get_state(), c.get_state())
        assert a.check_finite() == 0
        for s in (a, b, c):
            s.close()


def test_graph_rollout_stops_at_failing_residual_test():
    from mpm_snow_sim2d_amd._lib import MPM_E_STATE, MPMError
    from mpm_snow_sim2d_amd.solver import GPUSolver
    from mpm_snow_sim2d_amd.surrogate import ADVANCE_BALLISTIC, Surrogate
    P = snow_params(2, 64, 2e-4)
    rng = np.random.default_rng(5)
    c = np.array([0.5, 0.5])
    pts = jitter_lattice(c - 0.06, c + 0.06, 0.004, rng, lambda p: ((p - c) ** 2).sum(1) <= 0.06 ** 2)
    ps = _set(pts, (0.0, -3.0), 0.004 ** 2, 400 * 0.004 ** 2)     # towards the floor
    k = 8
    _flush(f"[rollout] gate, n={ps.n}")
    with Surrogate(64, 64, max_batch=1, capacity=0) as sur:
        # step by step: macro steps until the residual test fails
        b = GPUSolver(P, ps, Options(sort_every=4))
        before, taken = None, 0
        while True:
            st = b.get_state()
            try:
                sur.advance(b, k, ADVANCE_BALLISTIC, 1e-3)
            except MPMError as e:
                assert e.code == MPM_E_STATE
                before = st
                break
            taken += 1
            assert taken < 200
        b.close()
        assert taken >= 2
        # the recorded rollout stops at the same macro step
        a = GPUSolver(P, ps, Options(sort_every=4))
        with pytest.raises(MPMError) as err:
            sur.rollout_steps(a, k, taken + 10, ADVANCE_BALLISTIC, 1e-3)
        assert err.value.code == MPM_E_STATE
        assert f"macro step {taken + 1} " in str(err.value)
        # the particles hold the state before the failing step: the capture's float
        # atomics make the steps before it differ at round-off only
        from tests.helpers import state_errors
        err = state_errors(a.get_state(), before, float(P.dx))
        _flush(f"    rollout vs stepwise before the failing step {err}")
        for key, (mx, med) in err.items():
            assert mx < 1e-4 and med < 1e-6, (key, mx, med)
        a.close()
