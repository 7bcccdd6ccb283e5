"""The device SVDs (svd2 / svd3 in csrc/mpm_device.hpp) on the reference's edge
cases, through one substep of the C-ABI (GPU).

Particles of zero volume are skipped by P2G (p2g.glsl:42), so the grid stays empty,
every node velocity is zero and G2P sees F^_E = (I + dt C) F_E = F_E exactly
(g2p.glsl:74-75): the substep then applies the SVD, the clamp of the singular
values to [1 - theta_c, 1 + theta_s] and >= 1e-6, F_E' = U clamp(Sigma) V^T and
J_P' = J_P prod(sigma / clamp(sigma)) (g2p.glsl:80-93) to each particle's F_E.  The
cases are the branches the reference's SVDs take:

  2D (common.glsl:2-37)  the degenerate-angle guard (M^T M = c I: identity,
                          rotations, 2 I), reflections (det < 0: the last sigma
                          negative), rank-deficient and zero matrices, shear,
                          stretched, and an inverted general matrix
  3D (common.glsl:8-105) the Jacobi skip (off-diagonals below 1e-10, exactly
                          diagonal and near-diagonal), reflections on each axis
                          (the det U flip of the last column), rank-deficient,
                          rotated stretches, shear, and general matrices of both
                          determinant signs

Bound: the device's F_E' and J_P' may differ from the FP64 oracle by at most three
times what the FP32 oracle does on the same case, plus 2e-5 (FP32 resolution of
the reconstruction and of sqrt/reciprocal approximations), staged and fused
pipelines alike.
"""
import numpy as np
import pytest

from mpm_snow_sim2d_amd.params import Options, snow_params
from mpm_snow_sim2d_amd.scenes import ParticleSet
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _rot2(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s], [s, c]])


def _rot3(axis, a):
    c, s = np.cos(a), np.sin(a)
    i, j = [(1, 2), (0, 2), (0, 1)][axis]
    R = np.eye(3)
    R[i, i] = R[j, j] = c
    R[i, j], R[j, i] = -s, s
    return R


def _cases(dim):
    rng = np.random.default_rng(17)
    if dim == 2:
        cs = [np.eye(2), 2 * np.eye(2), _rot2(0.7), _rot2(-2.1), 0.98 * _rot2(1.3),
              np.diag([1.0, -1.0]), np.diag([-1.0, 1.0]), np.diag([-1.0, -1.0]), np.diag([1.1, -0.7]),
              np.array([[1.0, 0.0], [0.0, 0.0]]), np.zeros((2, 2)), np.array([[0.0, 1.0], [0.0, 0.0]]),
              np.array([[1.0, 0.4], [0.0, 1.0]]), np.diag([1.2, 0.8]), np.diag([1.004, 0.99]),
              np.array([[0.3, -1.1], [0.9, 0.2]]), np.array([[0.2, 1.3], [0.9, -0.4]])]
    else:
        cs = [np.eye(3), 2 * np.eye(3), np.diag([1.004, 0.98, 1.0]), np.diag([1.2, 0.9, 0.5]),
              np.eye(3) + 1e-11 * (np.ones((3, 3)) - np.eye(3)),
              np.eye(3) + 3e-9 * (np.ones((3, 3)) - np.eye(3)),
              np.diag([1.0, 1.0, -1.0]), np.diag([-1.0, 1.0, 1.0]), np.diag([1.0, -1.0, 1.0]),
              np.diag([-1.0, -1.0, -1.0]), np.diag([1.3, -0.6, 0.9]),
              np.diag([1.0, 1.0, 0.0]), np.zeros((3, 3)), np.outer([1.0, 2.0, 0.5], [0.3, -0.2, 1.0]),
              _rot3(2, 0.5) @ np.diag([1.01, 0.97, 1.0]) @ _rot3(0, -0.8),
              _rot3(1, 2.0) @ np.diag([1.1, 0.95, 0.9]) @ _rot3(2, 0.4).T,
              np.array([[1.0, 0.3, 0.0], [0.0, 1.0, 0.2], [0.0, 0.0, 1.0]])]
    for _ in range(4):
        M = np.eye(dim) + 0.3 * rng.normal(size=(dim, dim))
        cs.append(M if np.linalg.det(M) > 0 else -M)
        cs.append(-M if np.linalg.det(M) > 0 else M)
    return [np.asarray(c, np.float32) for c in cs]


def _scene(dim, cases):
    n = len(cases)
    P = snow_params(dim, 64 if dim == 2 else 32, 1e-4)
    x = np.full((dim, n), 0.5, np.float32)
    x[0] = 0.25 + 0.5 * (np.arange(n) + 0.37) / n          # interior, well inside the walls
    x[1] += 0.003 * np.sin(np.arange(n))
    ps = ParticleSet(x, np.zeros((dim, n), np.float32), np.full(n, 1e-4, np.float32), np.zeros(n, np.float32))
    F = np.stack([c.reshape(-1) for c in cases], axis=1).astype(np.float32)    # (dim*dim, n) row-major
    return P, ps, F


def _oracle(P, ps, F, dtype):
    st = orc.OracleState.from_particles(ps, dtype)
    st.Fe[:] = F.astype(dtype)
    orc.run(P, st, 1)
    s = orc.to_solver_state(P, st)
    return s["F"].astype(np.float64), s["Jp"].astype(np.float64), st.x.astype(np.float64)


@pytest.mark.parametrize("dim,fused", [(2, 0), (2, 1), (3, 0), (3, 1), (3, 2)])
def test_svd_edge_cases_match_oracle(dim, fused):
    """fused 2 (3D super-block items) takes elastic cases through the left stretch
    without an SVD and the others through svd3."""
    from mpm_snow_sim2d_amd.solver import GPUSolver
    cases = _cases(dim)
    P, ps, F = _scene(dim, cases)
    print(f"[svd] dim={dim} fused={fused}: {ps.n} cases", flush=True)
    with GPUSolver(P, ps, Options(fused=fused, sort_every=4), F=F) as s:
        s.Step(1)
        st = s.get_state()
    f64, j64, x64 = _oracle(P, ps, F, np.float64)
    f32, j32, _ = _oracle(P, ps, F, np.float32)
    dF_dev = np.abs(st["F"] - f64).max(axis=0)
    dF_32 = np.abs(f32 - f64).max(axis=0)
    dJ_dev = np.abs(st["Jp"] - j64)
    dJ_32 = np.abs(j32 - j64)
    for i in range(ps.n):
        assert dF_dev[i] <= 3 * dF_32[i] + 2e-5, (i, cases[i], dF_dev[i], dF_32[i])
        assert dJ_dev[i] <= 3 * dJ_32[i] + 2e-5 * max(1.0, abs(j64[i])), (i, cases[i], st["Jp"][i], j64[i])
    # the substep moved nothing (empty grid): positions reproduce themselves
    assert np.abs(st["x"] - x64).max() < 1e-6
    # and the cases exercise the clamp and the sign convention: some J_P' < 0
    # (reflections), some J_P' != 1 (plastic flow), some F_E' = F_E (elastic)
    assert (j64 < 0).any() and (np.abs(j64 - 1) > 1e-3).any()
    assert (np.abs(f64 - F).max(axis=0) < 1e-6).any()
