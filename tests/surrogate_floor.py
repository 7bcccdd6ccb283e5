"""Error floor of the surrogate's macro step with a perfect network (CPU).

Each macro step of k substeps takes its grid fields from the FP32 oracle run k
substeps ahead of the current state (the targets a perfect network would
predict), and the rollout is compared with the plain oracle run over the same
substeps.  The macro step is the NumPy restatement of k_macro_g2p
(mpm_snow_sim2d_amd/csrc/mpm_surrogate.hip) on the oracle's state (B = C / D^-1,
full F_P); the fields are restatements of the capture (k_capture, k_jacobi_*).

Forms:
  ballistic  the method without a network: the transfer of the current velocity
             plus the gravity impulse (MPM_ADVANCE_BALLISTIC, macro flags 0)
  transfer   the default surrogate: the transfer of the velocities k substeps
             ahead; trapezoid positions, F_E from T C' (flags 0, refinement 0)
  fit<J>     velocity and displacement fields, each the transfer refined by J
             Jacobi steps of the mass-weighted least-squares fit; positions and
             F_E from the displacement (flags 3, target refinement J)

  python tests/surrogate_floor.py [--seeds 5000,5001] [--substeps 2000] [--forms transfer,fit8]

prints the mean position error against the plain oracle per seed and form.
Test infrastructure (it runs the oracle); tests/test_surrogate_floor.py runs it
on a small scene.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402


class Image:
    """Quadratic B-spline transfers between particles and the h x w image of grid
    nodes [0, w) x [0, h) (nodes outside take no part, as on the device)."""

    def __init__(self, P, h: int, w: int):
        self.dx = float(P.dx)
        self.h, self.w = h, w

    def stencil(self, x):
        xg = np.asarray(x, np.float64) / self.dx
        base = np.floor(xg - 0.5).astype(np.int64)
        out = []
        for a in range(3):
            for b in range(3):
                i, j = base[0] + a, base[1] + b
                di, dj = xg[0] - i, xg[1] - j
                ai, aj = np.abs(di), np.abs(dj)
                wx = np.where(ai < 0.5, 0.75 - ai * ai, np.where(ai < 1.5, 0.5 * (1.5 - ai) ** 2, 0.0))
                wy = np.where(aj < 0.5, 0.75 - aj * aj, np.where(aj < 1.5, 0.5 * (1.5 - aj) ** 2, 0.0))
                ok = (i >= 0) & (i < self.w) & (j >= 0) & (j < self.h)
                q = np.where(ok, j * self.w + i, 0)
                # x_i - x_p in physical units
                dp = np.stack([-di * self.dx, -dj * self.dx])
                out.append((q, wx * wy, ok, dp))
        return out

    def scatter(self, st, vals):
        img = np.zeros((vals.shape[0], self.h * self.w))
        for q, wt, ok, _ in st:
            for c in range(vals.shape[0]):
                np.add.at(img[c], q[ok], (wt * vals[c])[ok])
        return img

    def gather(self, st, field):
        return sum(np.where(ok, wt, 0.0) * field[:, q] for q, wt, ok, _ in st)

    def transfer(self, st, m, y):
        """sum_p w_ip m_p y_p / m_i and the node mass m_i."""
        acc = self.scatter(st, np.vstack([m[None], m * y]))
        mass = acc[0]
        inv = np.where(mass > 0, 1.0 / np.where(mass > 0, mass, 1.0), 0.0)
        return acc[1:] * inv, mass, inv

    def fit(self, st, m, y, iters: int):
        """The transfer refined by `iters` steps t += P2G(m (y - G2P(t))) / m_i."""
        t, _, inv = self.transfer(st, m, y)
        for _ in range(iters):
            t = t + self.scatter(st, m * (y - self.gather(st, t))) * inv
        return t


def macro_step(P, img: Image, st: orc.OracleState, k: int, vfield, ufield=None):
    """k_macro_g2p on the oracle state: vfield the predicted velocity field (gravity
    included), ufield the predicted displacement (None: trapezoid positions and F_E
    from T C')."""
    dt = float(np.float32(P.DELTA_TIME))
    T = k * dt
    Dinv = 4.0 / img.dx ** 2
    th0, th1 = 1.0 - float(np.float32(P.THETA_C)), 1.0 + float(np.float32(P.THETA_S))
    g = float(np.float32(P.GRAVITY))
    ug = g * dt * dt * k * (k + 1) / 2
    x = st.x.astype(np.float64)
    v = st.v.astype(np.float64)
    n = x.shape[1]
    sten = img.stencil(x)
    nv = np.zeros((2, n))
    nB = np.zeros((2, 2, n))
    nu = np.zeros((2, n))
    nG = np.zeros((2, 2, n))
    for q, wt, ok, dp in sten:
        # nodes outside the image carry the particle's own ballistic motion
        vi = np.where(ok, vfield[:, q], v + np.array([[0.0], [g * T]]))
        nv += wt * vi
        nB += wt * vi[:, None, :] * dp[None, :, :]
        if ufield is not None:
            ui = np.where(ok, ufield[:, q], T * v + np.array([[0.0], [ug]]))
            nu += wt * ui
            nG += wt * ui[:, None, :] * dp[None, :, :]
    grad = Dinv * nG if ufield is not None else T * Dinv * nB
    M = np.eye(2)[None] + grad.transpose(2, 0, 1)
    F = st.Fe.astype(np.float64).reshape(2, 2, n).transpose(2, 0, 1)
    U, s, Vt = np.linalg.svd(M @ F)
    sc = np.maximum(np.clip(s, th0, th1), 1e-6)
    Fp = st.Fp.astype(np.float64).reshape(2, 2, n).transpose(2, 0, 1)
    Fp = (Vt.transpose(0, 2, 1) * (s / sc)[:, None, :]) @ Vt @ Fp
    st.Fe[:] = ((U * sc[:, None, :]) @ Vt).transpose(1, 2, 0).reshape(4, n)
    st.Fp[:] = Fp.transpose(1, 2, 0).reshape(4, n)
    st.x[:] = x + nu if ufield is not None else x + 0.5 * T * (v + nv)
    st.v[:] = nv
    st.B[:] = nB.reshape(4, n)


def rollout(P, ps, substeps: int, k: int, form: str, run=None) -> orc.OracleState:
    """Macro steps of `form` with perfect fields; run = the oracle's substep driver."""
    run = run or orc.run
    img = Image(P, P.HEIGHT, P.WIDTH)
    st = orc.OracleState.from_particles(ps, np.float32)
    m = st.mass.astype(np.float64)
    g = float(np.float32(P.GRAVITY))
    T = k * float(np.float32(P.DELTA_TIME))
    for _ in range(substeps // k):
        sten = img.stencil(st.x)
        if form == "ballistic":
            vf, _, _ = img.transfer(sten, m, st.v.astype(np.float64))
            vf[1] += g * T
            macro_step(P, img, st, k, vf)
            continue
        fut = st.copy()
        run(P, fut, k)
        if form == "transfer":
            vf, _, _ = img.transfer(sten, m, fut.v.astype(np.float64))
            macro_step(P, img, st, k, vf)
        elif form.startswith("fit"):
            iters = int(form[3:])
            y = np.vstack([fut.v.astype(np.float64), fut.x.astype(np.float64) - st.x.astype(np.float64)])
            t = img.fit(sten, m, y, iters)
            macro_step(P, img, st, k, t[:2], t[2:])
        else:
            raise ValueError(form)
    return st


def main() -> int:
    from mpm_snow_sim2d_amd.params import snow_params
    from mpm_snow_sim2d_amd.scenes import colliding_scene
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", default="5000,5001")
    ap.add_argument("--substeps", type=int, default=2000)
    ap.add_argument("--particles", type=int, default=60_000)
    ap.add_argument("--forms", default="ballistic,transfer,fit1,fit4,fit8")
    args = ap.parse_args()
    P = snow_params(2, 256, 5e-5)
    run = orc.run_omp      # the same substep on every host core
    for seed in (int(s) for s in args.seeds.split(",")):
        ps = colliding_scene(seed, args.particles)
        ref = orc.OracleState.from_particles(ps, np.float32)
        run(P, ref, args.substeps)
        for form in args.forms.split(","):
            t0 = time.perf_counter()
            st = rollout(P, ps, args.substeps, 8, form, run)
            err = float(np.linalg.norm(st.x.astype(np.float64) - ref.x, axis=0).mean())
            print(f"seed {seed} {form}: mean position error {err:.4e} m ({time.perf_counter() - t0:.0f} s)", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
