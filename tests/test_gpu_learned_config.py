"""The learned surrogate at its configured size (GPU): a 256^2 grid (dx = 1/256,
dt = 5e-5) and a colliding scene of ~60k particles, the shape the learned workload
of bench.py trains and rolls out on.

* capture: the node features (mass, velocity, mean J_P, mean log sigma of F_E) and
  the velocity / displacement targets k = 8 substeps later against a NumPy FP64
  P2G of the same particles (bounds of the small-scene test);
* network: one 256^2 image through the BF16 matrix-core U-Net against the FP64
  oracle (oracle/unet_ref.py) fed the same BF16-rounded weights (relative L2 1e-2);
* macro step: the ballistic macro step of the scene after its bodies have met
  (non-uniform velocity, C and F_E) against the NumPy restatement of k_macro_g2p
  (tests/surrogate_floor.py);
* rollout: 2000 substeps in macro steps that never fall back (the bodies stay
  clear of the walls), finite, and their positions against the plain solver's at
  the distance the macro step's own error allows (the ballistic method: a few cm).
"""
import numpy as np
import pytest

from mpm_snow_sim2d_amd.params import Options, snow_params
from mpm_snow_sim2d_amd.scenes import colliding_scene
from oracle import unet_ref as U
from tests.test_gpu_surrogate import _net_problem, _p2g_numpy, _rel

pytestmark = pytest.mark.gpu

CELLS, DT, NPART, K = 256, 5e-5, 60_000, 8


def _flush(msg):
    print(msg, flush=True)


def test_capture_at_config_size():
    from mpm_snow_sim2d_amd.solver import GPUSolver
    from mpm_snow_sim2d_amd.surrogate import Surrogate
    P = snow_params(2, CELLS, DT)
    ps = colliding_scene(5000, NPART)
    _flush(f"[learned-size] capture n={ps.n} on {CELLS}^2")
    with Surrogate(CELLS, CELLS, max_batch=1, capacity=2) as sur:
        s = GPUSolver(P, ps, Options(sort_every=8, tile_particles=256))
        s.Step(40)
        st0 = s.get_state()
        sur.record(s, 0, 1, 0, K)
        s.Step(K)
        st1 = s.get_state()
        sur.record(s, 0, 1, 1, K)
        data, sc = sur.get_sample(1)
        s.close()
    m = ps.mass.astype(np.float64)
    F = st0["F"].astype(np.float64)
    ls = 0.5 * np.log(np.abs(F[0] * F[3] - F[1] * F[2]))
    acc = _p2g_numpy(P, st0["x"], np.stack([st0["v"][0], st0["v"][1], st0["Jp"], ls]), m, CELLS, CELLS)
    y = np.concatenate([st1["v"].astype(np.float64), st1["x"].astype(np.float64) - st0["x"]])
    tgt = _p2g_numpy(P, st0["x"], y, m, CELLS, CELLS)
    mass = acc[0]
    occ = mass > 1e-6 * mass.max()
    assert np.allclose(data[0], mass, rtol=1e-5, atol=1e-6 * mass.max())
    for c in range(4):
        ref = acc[c + 1][occ] / mass[occ]
        tol = max(1e-4 * np.abs(ref).max(), 1e-6)
        assert np.abs(data[c + 1][occ] - ref).max() <= tol, c
    for c in range(4):
        ref = tgt[c + 1][occ] / mass[occ]
        assert np.abs(data[5 + c][occ] - ref).max() <= 2e-4 * np.abs(ref).max(), c
    assert abs(sc["particle_mass"] - m.sum()) <= 1e-5 * m.sum()
    assert sc["band_mass"] == 0.0


def test_network_forward_at_config_size():
    from mpm_snow_sim2d_amd.surrogate import Surrogate
    rng = np.random.default_rng(11)
    _flush(f"[learned-size] forward 1x{CELLS}x{CELLS}")
    with Surrogate(CELLS, CELLS, max_batch=1, seed=3) as sur:
        prm = sur.get_params()
        o, _ = U.layout()[0]["head.w"]
        prm[o:o + 64] = rng.normal(scale=0.5, size=64)
        sur.set_params(prm)
        p, bits, xr, wm, _ = _net_problem(rng, sur, 1, CELLS, CELLS)
        out = sur.forward(bits, wm)
    ref = U.forward(p, xr, wm.astype(np.float64))
    err = _rel(out, ref)
    _flush(f"    rel L2 {err:.3e}")
    assert err < 1e-2


def test_ballistic_macro_step_at_config_size():
    import surrogate_floor as SF
    from mpm_snow_sim2d_amd.solver import GPUSolver
    from mpm_snow_sim2d_amd.surrogate import ADVANCE_BALLISTIC, Surrogate
    from oracle import oracle as orc
    P = snow_params(2, CELLS, DT)
    ps = colliding_scene(5000, NPART)
    _flush(f"[learned-size] macro step n={ps.n}")
    with Surrogate(CELLS, CELLS, max_batch=1, capacity=0) as sur:
        s = GPUSolver(P, ps, Options(sort_every=8, tile_particles=256))
        s.Step(800)                 # the bodies have met and deform
        st0 = s.get_state()
        used, res = sur.advance(s, K, ADVANCE_BALLISTIC, 1e-3)
        st1 = s.get_state()
        s.close()
    assert used == ADVANCE_BALLISTIC
    n = ps.n
    dinv = 4.0 / float(P.dx) ** 2
    Fp = np.zeros((4, n))
    Fp[0] = st0["Jp"]
    Fp[3] = 1.0
    ost = orc.OracleState(st0["x"].astype(np.float64), st0["v"].astype(np.float64),
                          st0["C"].astype(np.float64) / dinv, st0["F"].astype(np.float64), Fp,
                          ps.mass.astype(np.float64), ps.vol.astype(np.float64))
    img = SF.Image(P, CELLS, CELLS)
    sten = img.stencil(ost.x)
    dt = float(np.float32(P.DELTA_TIME))
    T, g = K * dt, float(np.float32(P.GRAVITY))
    vf, _, _ = img.transfer(sten, ost.mass, ost.v)
    vf[1] += g * T
    SF.macro_step(P, img, ost, K, vf, None)
    vscale = np.abs(ost.v).max()
    assert np.abs(st1["x"] - ost.x).max() < 1e-6
    assert np.abs(st1["v"] - ost.v).max() < 1e-4 * vscale
    C = dinv * ost.B
    assert np.abs(st1["C"] - C).max() < 1e-3 * max(np.abs(C).max(), 1.0)
    assert np.abs(st1["F"] - ost.Fe).max() < 1e-5
    Jp = ost.Fp[0] * ost.Fp[3] - ost.Fp[1] * ost.Fp[2]
    assert np.abs(st1["Jp"] - Jp).max() < 1e-5
    assert np.abs(st1["F"] - st0["F"]).max() > 1e-4


def test_ballistic_rollout_at_config_size_never_falls_back():
    from mpm_snow_sim2d_amd.surrogate import ADVANCE_BALLISTIC, ADVANCE_PLAIN, Surrogate, rollout
    P = snow_params(2, CELLS, DT)
    ps = colliding_scene(5000, NPART)
    nsub = 2000
    _flush(f"[learned-size] rollout n={ps.n} {nsub} substeps")
    ref = rollout(None, P, ps, nsub, K, ADVANCE_PLAIN)
    with Surrogate(CELLS, CELLS, max_batch=1, capacity=0) as sur:
        r = rollout(sur, P, ps, nsub, K, ADVANCE_BALLISTIC, tol=1e-3)
    assert r["counts"][ADVANCE_BALLISTIC] == nsub // K and r["counts"][ADVANCE_PLAIN] == 0
    assert r["nonfinite"] == 0 and ref["nonfinite"] == 0
    err = float(np.linalg.norm(r["x"] - ref["x"], axis=1).mean())
    _flush(f"    mean position error of the ballistic method {err:.4f} m")
    # the method without a network drifts by centimetres over the rollout (the bodies
    # collide), not by the domain
    assert 1e-3 < err < 0.1
