"""The surrogate macro step's error floor with a perfect network (CPU, through the
oracle; the restatement lives in tests/surrogate_floor.py).

* The NumPy restatement of the macro step reproduces free fall in closed form (the
  same check the device macro step passes), so it is pinned independently.
* On a small colliding scene, with every macro step fed the oracle's own future
  (the targets a perfect network would predict): the default form (transfer of
  the future velocities, trapezoid positions, F_E from T C') stays far closer to the
  plain oracle than the ballistic method, and the least-squares velocity +
  displacement form closer still.  At the learned workload's size
  (`python tests/surrogate_floor.py`) the same ordering holds with the numbers
  recorded in the design notes.
"""
import numpy as np
import pytest

from mpm_snow_sim2d_amd.params import snow_params
from mpm_snow_sim2d_amd.scenes import _set, jitter_lattice

import surrogate_floor as SF


def _two_blocks(speed=3.0, h=0.005):
    rng = np.random.default_rng(3)
    out = None
    for cx, vx in ((0.40, speed), (0.60, -speed)):
        c = np.array([cx, 0.5])
        pts = jitter_lattice(c - 0.08, c + 0.08, h, rng, lambda p, c=c: (np.abs(p - c) <= 0.08).all(1))
        ps = _set(pts, (vx, 0.0), h * h, 400 * h * h)
        out = ps if out is None else out.concat(ps)
    return out


def test_macro_restatement_free_fall(orc):
    P = snow_params(2, 64, 2e-4)
    rng = np.random.default_rng(4)
    c = np.array([0.5, 0.55])
    pts = jitter_lattice(c - 0.08, c + 0.08, 0.004, rng, lambda p: ((p - c) ** 2).sum(1) <= 0.08 ** 2)
    ps = _set(pts, (0.7, -0.3), 0.004 ** 2, 400 * 0.004 ** 2)
    k = 8
    for displacement in (False, True):
        st = orc.OracleState.from_particles(ps, np.float32)
        x0, v0 = st.x.astype(np.float64), st.v.astype(np.float64)
        img = SF.Image(P, 64, 64)
        sten = img.stencil(st.x)
        m = st.mass.astype(np.float64)
        dt = float(np.float32(P.DELTA_TIME))
        T, g = k * dt, float(np.float32(P.GRAVITY))
        vf, _, _ = img.transfer(sten, m, v0)
        vf[1] += g * T
        uf = None
        if displacement:
            uf, _, _ = img.transfer(sten, m, T * v0)
            uf[1] += g * dt * dt * k * (k + 1) / 2
        SF.macro_step(P, img, st, k, vf, uf)
        fall = dt * dt * k * (k + 1) / 2 if displacement else 0.5 * T * T
        assert np.abs(st.v - (v0 + np.array([[0.0], [g * T]]))).max() < 1e-6
        assert np.abs(st.x - (x0 + T * v0 + np.array([[0.0], [g * fall]]))).max() < 1e-6
        assert np.abs(st.Fe - np.array([[1.0], [0.0], [0.0], [1.0]])).max() < 1e-6


@pytest.mark.timeout(600)
def test_perfect_network_floor_ordering(orc):
    P = snow_params(2, 64, 2e-4)
    ps = _two_blocks()
    n = 160                      # the blocks meet after ~40 substeps and compact
    ref = orc.OracleState.from_particles(ps, np.float32)
    orc.run(P, ref, n)
    err = {}
    for form in ("ballistic", "transfer", "fit8"):
        st = SF.rollout(P, ps, n, 8, form)
        err[form] = float(np.linalg.norm(st.x.astype(np.float64) - ref.x, axis=0).mean())
    print(err)
    assert err["transfer"] < 0.5 * err["ballistic"]
    assert err["fit8"] < err["transfer"]
