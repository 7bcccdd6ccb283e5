"""Pins the oracle's restatement of the CPU path's scheme (src/cpu: APIC transfer
with an explicit grid force) and the stress terms of both schemes through closed-form
moments of one particle's grid contributions.  CPU only.

Moments used (w, grad w from one particle at x_p, all stencil nodes inside):
  sum_i grad w_i = 0,  sum_i x_i (grad w_i)^T = I,  sum_i w_i (x_i - x_p)(x_i - x_p)^T = D
so for the CPU scheme  sum_i f_i = 0 and sum_i f_i x_i^T = A_p = V tau, and for the
GPU (MLS) scheme  sum_i mv_i (x_i - x_p)^T = m C D - dt V tau.
"""
import numpy as np
import pytest

from mpm_snow_sim2d_amd.params import SCHEME_APIC_FORCE, reference_params, snow_params
from mpm_snow_sim2d_amd.scenes import ParticleSet


def _one(dim, P, F, Jp=1.0, vel=None, Bm=None, pos=None):
    pos = np.asarray(pos if pos is not None else [60.3, 40.7, 41.2][:dim], np.float64) * (P.dx if P.dx != 1 else 1)
    ps = ParticleSet(pos[:, None].copy(), np.zeros((dim, 1)) if vel is None else np.asarray(vel, float)[:, None],
                     np.array([2.5]), np.array([0.7]))
    return ps


def _node_coords(P):
    d = P.dim
    W1, H1 = P.WIDTH + 1, P.HEIGHT + 1
    ids = np.arange(P.num_nodes)
    if d == 2:
        return np.stack([ids % W1, ids // W1]).astype(float) * P.dx
    return np.stack([ids % W1, (ids // W1) % H1, ids // (W1 * H1)]).astype(float) * P.dx


@pytest.mark.parametrize("dim,order", [(2, 3), (3, 3), (2, 2), (3, 2)])
def test_cpu_scheme_force_moments(orc, dim, order):
    P = reference_params(dim) if order == 3 else snow_params(dim, 32, 1e-4)
    P.scheme = SCHEME_APIC_FORCE
    if dim == 3 and order == 3:
        P.WIDTH, P.HEIGHT, P.DEPTH = 80, 60, 60
    rng = np.random.default_rng(dim + order)
    F = np.eye(dim) + rng.normal(scale=0.03, size=(dim, dim))
    ps = _one(dim, P, F, pos=[20.3, 17.6, 18.9][:dim] if order == 3 else [15.3, 14.6, 16.2][:dim])
    st = orc.OracleState.from_particles(ps, np.float64)
    st.Fe[:, 0] = F.reshape(-1)
    Jp = 0.97
    st.Fp[:, 0] = np.diag([Jp] + [1.0] * (dim - 1)).reshape(-1)
    gm, gmom, gf = orc.p2g(P, st, with_force=True)
    xi = _node_coords(P)
    tau = orc.stress(P, F, Jp)
    A = st.vol[0] * tau
    assert np.abs(gf.sum(1)).max() < 1e-9 * np.abs(A).max()
    assert np.allclose(gf @ xi.T, A, rtol=1e-9, atol=1e-9 * np.abs(A).max())


@pytest.mark.parametrize("dim,order", [(2, 3), (3, 2)])
def test_gpu_scheme_stress_moment(orc, dim, order):
    """MLS: the stress enters the momentum with first moment -dt V tau."""
    P = reference_params(dim) if order == 3 else snow_params(dim, 32, 1e-4)
    if dim == 3 and order == 3:
        P.WIDTH, P.HEIGHT, P.DEPTH = 80, 60, 60
    rng = np.random.default_rng(7)
    F = np.eye(dim) + rng.normal(scale=0.03, size=(dim, dim))
    ps = _one(dim, P, F, pos=[20.3, 17.6, 18.9][:dim] if order == 3 else [15.3, 14.6, 16.2][:dim])
    st = orc.OracleState.from_particles(ps, np.float64)
    st.Fe[:, 0] = F.reshape(-1)
    gm, gmom = orc.p2g(P, st)
    xi = _node_coords(P)
    rel = xi - st.x[:, :1]
    mom1 = gmom @ rel.T
    tau = orc.stress(P, F, 1.0)
    dt = np.float64(np.float32(P.DELTA_TIME))
    assert np.allclose(mom1, -dt * st.vol[0] * tau, rtol=1e-8, atol=1e-10 * np.abs(tau).max())


def test_cpu_scheme_free_fall_and_affine(orc):
    """Stress-free motion is identical in both schemes: free fall, and an affine
    velocity field whose gradient the CPU scheme recovers as grad v."""
    P = reference_params(2)
    P.scheme = SCHEME_APIC_FORCE
    P.GRAVITY = 0.0
    axes = [np.arange(60.25, 68, 0.5), np.arange(40.25, 46, 0.5)]
    pts = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 2).T
    n = pts.shape[1]
    G = np.array([[0.05, -0.4], [0.4, 0.02]])
    c = np.array([64.0, 43.0])
    ps = ParticleSet(pts.copy(), (G @ (pts - c[:, None])), np.full(n, 1.0), np.full(n, 0.25))
    st = orc.OracleState.from_particles(ps, np.float64)
    st.B = np.repeat((G / 3.0).reshape(-1)[:, None], n, 1)
    x0 = st.x.copy()
    st.Fe[:] = np.eye(2).reshape(-1)[:, None]
    orc.run(P, st, 1)
    dt = np.float64(np.float32(P.DELTA_TIME))
    field = G @ (x0 - c[:, None])
    assert np.allclose(st.v, field, atol=1e-10)
    # F_hat = (I + dt grad v) F with grad v = G, no plasticity for this small step
    Fe = st.Fe[:, 0].reshape(2, 2)
    assert np.allclose(Fe @ st.Fp[:, 0].reshape(2, 2), np.eye(2) + dt * G, atol=1e-9)


def test_cpu_scheme_conserves_momentum_without_walls(orc):
    P = snow_params(2, 64, 1e-4)
    P.scheme = SCHEME_APIC_FORCE
    P.GRAVITY = 0.0
    rng = np.random.default_rng(3)
    n = 400
    pts = 0.5 + rng.uniform(-0.06, 0.06, size=(2, n))
    ps = ParticleSet(pts, rng.normal(size=(2, n)), np.full(n, 1e-3), np.full(n, 1e-5))
    st = orc.OracleState.from_particles(ps, np.float64)
    st.Fe = st.Fe + rng.normal(scale=0.01, size=st.Fe.shape)
    p0 = (st.mass * st.v).sum(1)
    orc.run(P, st, 3)
    assert np.allclose((st.mass * st.v).sum(1), p0, rtol=1e-9, atol=1e-12)


def test_cpu_scheme_nonneg_svd_reflection(orc):
    """For an inverted F the CPU scheme keeps the reflection (non-negative sigma,
    no sign fix) where the GPU scheme flips the last sigma: F_E differs."""
    P = reference_params(2)
    P.GRAVITY = 0.0
    ps = ParticleSet(np.array([[60.3], [40.6]]), np.zeros((2, 1)), np.array([1.0]), np.array([0.25]))
    Finv = np.array([[1.0, 0.0], [0.0, -0.99]])
    outs = []
    for scheme in (0, SCHEME_APIC_FORCE):
        P.scheme = scheme
        st = orc.OracleState.from_particles(ps, np.float64)
        st.Fe[:, 0] = Finv.reshape(-1)
        st.vol[:] = 0.0          # no stress: isolate the plasticity step
        orc.run(P, st, 1)
        outs.append(st.Fe[:, 0].reshape(2, 2))
    assert np.linalg.det(outs[1]) < 0 < np.linalg.det(outs[0])
