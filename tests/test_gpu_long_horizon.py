"""Long-horizon and contact-regime parity of the plain solver (GPU).

* configs[0] at the benchmark's size (two snowballs, 19,979 particles, 128^2,
  dt = 1e-4, graph replay with the autotuner's usual options) for its full 1000
  substeps, through the collision of the two bodies: centre of mass, linear
  momentum, kinetic energy, sum of |m v| and mean J_P against the FP64 oracle at
  250, 500, 750 and 1000 substeps, within 0.1% of each aggregate's scale (the
  configuration's target is 1%); the FP32 oracle is checked the same way, which
  shows the aggregates are well conditioned through the collision (it meets the
  bound by two orders of magnitude).
* configs[2] at full size (16M particles in four spheres, 512^3, dt = 2.5e-5) for
  400 substeps, through the first contact of two spheres (~157 substeps): no
  particle reaches a wall in that time, so MLS-MPM's transfers conserve linear
  momentum and the total momentum changes by exactly the gravity impulse M g t; in
  FP32 it may drift by at most one unit of FP32 resolution (2^-23) of sum |m v| per
  substep (measured: about a fifth of that); the state stays finite, and the
  contact deforms the material plastically.
* configs[1] at full size (the 2,000,376-particle falling cube, 256^3, dt = 5e-5)
  for 1000 substeps of free fall: total momentum M g t (to 2^-23 per substep, as
  above), every particle's velocity
  within FP32 noise of g t, no plastic flow (J_P stays exactly 1: the singular
  values never leave the elastic band), F_E within elastic noise of I.
"""
import numpy as np
import pytest

from mpm_snow_sim2d_amd.params import Options, snow_params
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _aggregates(x, v, Jp, m):
    x = np.asarray(x, np.float64)
    v = np.asarray(v, np.float64)
    M = m.sum()
    return {"com": (m * x).sum(1) / M, "mom": (m * v).sum(1), "ke": 0.5 * float((m * (v * v).sum(0)).sum()),
            "mabs": float((m * np.linalg.norm(v, axis=0)).sum()), "jp": float(np.asarray(Jp, np.float64).mean())}


def _agg_errors(a, b, extent):
    """|a - b| per aggregate in its scale: centre of mass in body extents, momentum
    in sum |m v|, kinetic energy, sum |m v| and mean J_P relative."""
    return {"com": float(np.abs(a["com"] - b["com"]).max() / extent),
            "mom": float(np.abs(a["mom"] - b["mom"]).max() / b["mabs"]),
            "ke": abs(a["ke"] - b["ke"]) / b["ke"], "mabs": abs(a["mabs"] - b["mabs"]) / b["mabs"],
            "jp": abs(a["jp"] - b["jp"]) / abs(b["jp"])}


@pytest.mark.timeout(900)
def test_snowballs2d_aggregates_through_collision():
    from mpm_snow_sim2d_amd.scenes import snowball_disks
    from mpm_snow_sim2d_amd.solver import GPUSolver
    P = snow_params(2, 128, 1e-4)
    ps = snowball_disks(seed=0)
    assert ps.n == 19_979
    m = ps.mass.astype(np.float64)
    extent = 0.16                                     # a snowball's diameter
    checkpoints = [250, 500, 750, 1000]
    dev = []
    _flush(f"[long] snowballs2d n={ps.n}: device")
    with GPUSolver(P, ps, Options(sort_every=64, tile_particles=128)) as s:
        s.set_graph(True)
        done = 0
        for k in checkpoints:
            s.Step(k - done)
            done = k
            st = s.get_state()
            dev.append(_aggregates(st["x"], st["v"], st["Jp"], m))
        assert s.check_finite() == 0
    refs = {}
    for dt in (np.float64, np.float32):
        _flush(f"[long] oracle {dt.__name__}")
        o = orc.OracleState.from_particles(ps, dt)
        out, done = [], 0
        for k in checkpoints:
            orc.run(P, o, k - done)
            done = k
            so = orc.to_solver_state(P, o)
            out.append(_aggregates(so["x"], so["v"], so["Jp"], m))
        refs[dt.__name__] = out
    for i, k in enumerate(checkpoints):
        e_dev = _agg_errors(dev[i], refs["float64"][i], extent)
        e_32 = _agg_errors(refs["float32"][i], refs["float64"][i], extent)
        _flush(f"    {k:5d}: device {({a: f'{b:.1e}' for a, b in e_dev.items()})}  "
               f"FP32 oracle {({a: f'{b:.1e}' for a, b in e_32.items()})}")
        for key in e_dev:
            assert e_dev[key] < 1e-3, (k, key, e_dev[key])
            assert e_32[key] < 1e-3, (k, key, e_32[key])
    # the bodies did collide: the kinetic energy fell and J_P moved away from 1
    assert refs["float64"][-1]["ke"] < 0.5 * refs["float64"][0]["ke"]
    assert abs(refs["float64"][-1]["jp"] - 1.0) > 0.05


@pytest.mark.timeout(900)
@pytest.mark.parametrize("fused", [1, 2, 3])
def test_large_spheres_momentum_through_contact(fused):
    from mpm_snow_sim2d_amd.scenes import LARGE_SEED, colliding_spheres
    from mpm_snow_sim2d_amd.solver import GPUSolver
    P = snow_params(3, 512, 2.5e-5)
    ps = colliding_spheres(seed=LARGE_SEED)
    m = ps.mass.astype(np.float64)
    M = m.sum()
    v0 = ps.v.astype(np.float64)
    p0 = (m * v0).sum(1)
    scale = float((m * np.linalg.norm(v0, axis=0)).sum())
    dt = float(np.float32(P.DELTA_TIME))
    g = float(np.float32(P.GRAVITY))
    wall = P.BORDER_MARGIN * float(P.dx)
    _flush(f"[long] large spheres n={ps.n}: 400 substeps through first contact, fused={fused}")
    with GPUSolver(P, ps, Options(sort_every=32, fused=fused, tile_particles=512 if fused == 1 else 0)) as s:
        done = 0
        for k in (100, 200, 300, 400):
            s.Step(k - done)
            done = k
            st = s.get_state()
            x = st["x"]
            assert ((x > wall) & (x < 1.0 - wall)).all()          # nothing reached a wall
            p = (m * st["v"].astype(np.float64)).sum(1)
            expect = p0 + np.array([0.0, M * g * k * dt, 0.0])
            err = float(np.abs(p - expect).max() / scale)
            dF = float(np.abs(st["F"] - np.eye(3).reshape(9, 1)).max())
            plastic = int((np.abs(st["Jp"] - 1.0) > 1e-3).sum())
            _flush(f"    {k}: momentum error {err:.2e} of sum|m v|, max|F_E - I| {dF:.2e}, "
                   f"{plastic} particles with |J_P - 1| > 1e-3")
            assert err < k * 2.0 ** -23, (k, err)
        assert s.check_finite() == 0
    # the spheres met: the contact deformed material plastically
    assert plastic > 1000 and dF > 1e-3


@pytest.mark.timeout(900)
@pytest.mark.parametrize("fused", [1, 2])
def test_falling_cube_free_fall_invariants_1000_substeps(fused):
    from mpm_snow_sim2d_amd.scenes import falling_cube
    from mpm_snow_sim2d_amd.solver import GPUSolver
    P = snow_params(3, 256, 5e-5)
    ps = falling_cube(seed=0, n_side=126)
    m = ps.mass.astype(np.float64)
    M = m.sum()
    dt = float(np.float32(P.DELTA_TIME))
    g = float(np.float32(P.GRAVITY))
    _flush(f"[long] falling cube n={ps.n}: 1000 substeps of free fall, fused={fused}")
    with GPUSolver(P, ps, Options(sort_every=32, fused=fused, tile_particles=512 if fused == 1 else 0)) as s:
        done = 0
        for k in (250, 500, 1000):
            s.Step(k - done)
            done = k
            st = s.get_state()
            vy = g * k * dt
            p = (m * st["v"].astype(np.float64)).sum(1)
            dv = float(np.abs(st["v"] - np.array([0.0, vy, 0.0])[:, None]).max())
            dF = float(np.abs(st["F"] - np.eye(3).reshape(9, 1)).max())
            _flush(f"    {k}: momentum {p / M} vs g t {vy:.6f}; max |v - g t| {dv:.2e}; max|F_E - I| {dF:.2e}")
            assert np.abs(p - np.array([0.0, M * vy, 0.0])).max() <= k * 2.0 ** -23 * M * abs(vy)
            assert dv <= 1e-3 * abs(vy)
            assert dF < 1e-4
            assert np.all(st["Jp"] == 1.0)
        assert s.check_finite() == 0
