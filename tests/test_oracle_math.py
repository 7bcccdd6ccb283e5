"""Pins the oracle's device math (B-spline, SVD, stress) against closed forms,
invariants and LAPACK.  CPU only."""
import numpy as np
import pytest

from mpm_snow_sim2d_amd.params import reference_params, snow_params


# ----------------------------------------------------------------- B-spline (shaders/2d/common.glsl:43-82)
@pytest.mark.parametrize("order,reach,Dval", [(3, 2, 1.0 / 3.0), (2, 1.5, 0.25)])
def test_spline_moments(orc, order, reach, Dval):
    """Partition of unity, linear reproduction and the second moment D (cubic 1/3,
    quadratic 1/4 in grid units) at random particle offsets."""
    rng = np.random.default_rng(1)
    for fx in rng.uniform(0, 1, 50):
        base = -1 if order == 3 else (0 if fx >= 0.5 else -1)
        nodes = np.arange(base, base + (4 if order == 3 else 3))
        d = fx - nodes
        w = np.array([orc.weight(float(t), order) for t in d])
        assert abs(w.sum() - 1.0) < 1e-12
        assert abs((w * (nodes - fx)).sum()) < 1e-12
        assert abs((w * (nodes - fx) ** 2).sum() - Dval) < 1e-12


def test_spline_known_values(orc):
    cub = {0.0: 2 / 3, 0.5: 23 / 48, 1.0: 1 / 6, 1.5: 1 / 48, 2.0: 0.0, 2.5: 0.0}
    for x, want in cub.items():
        assert abs(orc.weight(x, 3) - want) < 1e-15
        assert abs(orc.weight(-x, 3) - want) < 1e-15
    quad = {0.0: 0.75, 0.5: 0.5, 1.0: 0.125, 1.5: 0.0, 0.25: 0.6875}
    for x, want in quad.items():
        assert abs(orc.weight(x, 2) - want) < 1e-15


@pytest.mark.parametrize("order", [2, 3])
def test_spline_gradient_matches_finite_difference(orc, order):
    h = 1e-6
    for x in np.linspace(-1.9, 1.9, 77):
        if order == 2 and min(abs(abs(x) - 0.5), abs(abs(x) - 1.5)) < 1e-3:
            continue
        fd = (orc.weight(x + h, order) - orc.weight(x - h, order)) / (2 * h)
        assert abs(orc.dweight(x, order) - fd) < 1e-6


def test_spline_f32_matches_f64(orc):
    for x in np.linspace(-2.2, 2.2, 101):
        for o in (2, 3):
            assert abs(orc.weight(float(np.float32(x)), o, np.float32) - orc.weight(float(np.float32(x)), o)) < 2e-7


# ----------------------------------------------------------------- SVD (2d/common.glsl:2-37, 3d/common.glsl:8-105)
def _check_svd(orc, M, tol):
    U, s, V = orc.svd(M)
    d = M.shape[0]
    assert np.abs(U @ np.diag(s) @ V.T - M).max() <= tol * max(1.0, np.abs(M).max())
    assert np.abs(V.T @ V - np.eye(d)).max() <= tol
    assert np.linalg.det(U) > 0
    assert np.linalg.det(V) > 0            # products of rotations
    # |sigma| are the singular values (LAPACK), in whatever order Jacobi leaves them
    assert np.allclose(np.sort(np.abs(s)), np.sort(np.linalg.svd(M.astype(np.float64))[1]), atol=tol * 10, rtol=tol * 10)
    # the sign convention: only the last sigma may be negative, and only if det M < 0
    assert (s[:-1] >= 0).all()
    if np.linalg.det(M.astype(np.float64)) < -1e-6:
        assert s[-1] < 0
    return U, s, V


@pytest.mark.parametrize("d", [2, 3])
def test_svd_random(orc, d):
    rng = np.random.default_rng(d)
    for _ in range(200):
        M = np.eye(d) + rng.normal(scale=0.3, size=(d, d))
        _check_svd(orc, M, 1e-12)
        _check_svd(orc, M.astype(np.float32), 2e-5)


@pytest.mark.parametrize("d", [2, 3])
def test_svd_special_cases(orc, d):
    I = np.eye(d)
    U, s, V = _check_svd(orc, I, 1e-14)
    assert np.allclose(s, 1)
    D = np.diag([1.5, 0.7, 1.1][:d])
    U, s, V = _check_svd(orc, D, 1e-14)
    assert np.allclose(np.sort(s), np.sort(np.diag(D)))
    th = 0.7
    R2 = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    R = R2 if d == 2 else np.block([[R2, np.zeros((2, 1))], [np.zeros((1, 2)), np.ones((1, 1))]])
    U, s, V = _check_svd(orc, R, 1e-13)
    assert np.allclose(s, 1)
    assert np.allclose(U @ V.T, R)        # polar rotation of a rotation is itself
    # reflection: the last singular value comes out negative (2d/common.glsl:32-36)
    Fl = np.eye(d)
    Fl[-1, -1] = -1
    U, s, V = _check_svd(orc, Fl, 1e-14)
    assert np.isclose(s[-1], -1) and np.allclose(s[:-1], 1)
    # rank deficient: sigma clamps at sqrt(1e-12) = 1e-6
    Z = np.zeros((d, d))
    Z[0, 0] = 2.0
    U, s, V = orc.svd(Z)
    assert np.isclose(np.abs(s).max(), 2.0)
    assert np.isclose(np.sort(np.abs(s))[0], 1e-6)


def test_svd2_degenerate_angle_branch(orc):
    """b = 0 and a = d takes theta = 0 (2d/common.glsl:13-15): V = I."""
    M = np.array([[0.0, 2.0], [2.0, 0.0]])          # M^T M = 4 I
    U, s, V = orc.svd(M)
    assert np.array_equal(V, np.eye(2))
    assert np.allclose(U @ np.diag(s) @ V.T, M)


def test_svd3_jacobi_skip_branch(orc):
    """Off-diagonals below 1e-10 skip the rotation (3d/common.glsl:10-15): V = I."""
    M = np.diag([1.3, 0.8, 1.05])
    U, s, V = orc.svd(M)
    assert np.array_equal(V, np.eye(3))
    assert np.array_equal(s, np.diag(M))


@pytest.mark.parametrize("d", [2, 3])
def test_svd_f32_close_to_f64(orc, d):
    rng = np.random.default_rng(10 + d)
    for _ in range(100):
        M = np.eye(d) + rng.normal(scale=0.05, size=(d, d))
        U64, s64, V64 = orc.svd(M)
        U32, s32, V32 = orc.svd(M.astype(np.float32))
        R64 = U64 @ V64.T
        R32 = U32.astype(np.float64) @ V32.T.astype(np.float64)
        assert np.abs(R64 - R32).max() < 5e-6
        assert np.abs(np.sort(s64) - np.sort(s32)).max() < 5e-6


# ----------------------------------------------------------------- stress (2d/p2g.glsl:44-59)
@pytest.mark.parametrize("dim", [2, 3])
def test_stress_closed_forms(orc, dim):
    P = reference_params(dim)
    I = np.eye(dim)
    assert np.abs(orc.stress(P, I, 1.0)).max() == 0.0
    th = 0.4
    R2 = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    R = R2 if dim == 2 else np.block([[R2, np.zeros((2, 1))], [np.zeros((1, 2)), np.ones((1, 1))]])
    assert np.abs(orc.stress(P, R, 1.0)).max() < 1e-9 * P.MU
    # F = s I: tau_ii = 2 mu (s-1) s + lambda (J-1) J with J = s^dim
    mu, la = np.float64(P.MU), np.float64(P.LAMBDA)
    for s_ in (0.97, 1.0, 1.004):
        tau = orc.stress(P, s_ * I, 1.0)
        J = s_ ** dim
        want = 2 * mu * (s_ - 1) * s_ + la * (J - 1) * J
        assert np.allclose(np.diag(tau), want, rtol=1e-10, atol=1e-9)
        assert np.abs(tau - np.diag(np.diag(tau))).max() < 1e-9 * mu
    # hardening exp(XI (1 - Jp)) scales the whole stress
    F = np.diag([1.01, 0.98, 1.0][:dim])
    t1 = orc.stress(P, F, 1.0)
    t2 = orc.stress(P, F, 0.95)
    assert np.allclose(t2, t1 * np.exp(P.XI * 0.05), rtol=1e-10)


def test_stress_symmetric_and_rotation_covariant(orc):
    P = snow_params(3, 64, 1e-4)
    rng = np.random.default_rng(3)
    for _ in range(20):
        F = np.eye(3) + rng.normal(scale=0.02, size=(3, 3))
        tau = orc.stress(P, F, 1.0)
        assert np.abs(tau - tau.T).max() < 1e-8 * np.abs(tau).max()
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        if np.linalg.det(q) < 0:
            q[:, 0] *= -1
        tq = orc.stress(P, q @ F, 1.0)          # tau(QF) = Q tau(F) Q^T
        assert np.allclose(tq, q @ tau @ q.T, rtol=1e-8, atol=1e-6)
