"""The oracle's G2P takes positions from the collision velocity and velocity / B
from the friction velocity (g2p.glsl:55-65), pinned in closed form (CPU).

With the two grid fields affine and different, v_col(x) = a_c + G_c x and
v_fric(x) = a_f + G_f x, the B-spline identities give for every particle whose
stencil lies inside the grid (partition of unity, linear reproduction, second
moment D = dx^2/4 quadratic, dx^2/3 cubic):

  v_p' = a_f + G_f x_p                            (velocity: friction field)
  B_p' = G_f D,   C_p' = D^-1 B_p' = G_f           (affine matrix: friction field)
  x_p' = x_p + dt (a_c + G_c x_p)                  (position: collision field)
  F_E' = clamp(I + dt G_f) F_E                     (deformation: C, friction field)

so exchanging the two fields anywhere in G2P changes v, C or x by
dt (a_c - a_f) or (G_c - G_f), which the bounds below resolve many times over.
"""
import numpy as np
import pytest

from mpm_snow_sim2d_amd.params import snow_params, reference_params


def _fields(P, dim, a, G):
    """Affine node field a + G x_i on the oracle's node numbering
    (node id = (k (H+1) + j) (W+1) + i, shaders/3d/p2g.glsl:75), (dim, nodes)."""
    W1, H1 = P.WIDTH + 1, P.HEIGHT + 1
    D1 = P.DEPTH + 1 if dim == 3 else 1
    k, j, i = np.meshgrid(np.arange(D1), np.arange(H1), np.arange(W1), indexing="ij")
    xi = np.stack([i.ravel(), j.ravel(), k.ravel()][:dim]) * float(P.dx)
    return (np.asarray(a)[:, None] + np.asarray(G) @ xi).astype(np.float64)


@pytest.mark.parametrize("dim,order", [(2, 2), (2, 3), (3, 2), (3, 3)])
def test_g2p_positions_from_collision_velocity_rest_from_friction(orc, dim, order):
    if order == 2:
        P = snow_params(dim, 32 if dim == 3 else 64, 1e-4)
    else:
        P = reference_params(dim)
        P.WIDTH, P.HEIGHT, P.DEPTH = 40, 30, 30
    rng = np.random.default_rng(21 + dim + order)
    n = 40
    dx = float(P.dx)
    lo, hi = 6 * dx, np.array([P.WIDTH, P.HEIGHT, P.DEPTH][:dim]) * dx - 6 * dx
    x = lo + rng.uniform(size=(dim, n)) * (hi - lo)[:, None]
    a_c, a_f = rng.normal(size=dim), rng.normal(size=dim)
    G_c, G_f = 0.5 * rng.normal(size=(dim, dim)), 0.5 * rng.normal(size=(dim, dim))
    vc = _fields(P, dim, a_c, G_c)
    vf = _fields(P, dim, a_f, G_f)
    eye = np.eye(dim).reshape(-1)
    st = orc.OracleState(x.copy(), np.zeros((dim, n)), np.zeros((dim * dim, n)),
                         np.repeat(eye[:, None], n, 1), np.repeat(eye[:, None], n, 1), np.full(n, 1e-3),
                         np.full(n, 1e-6))
    orc.g2p(P, st, vc, vf)
    dt = float(P.DELTA_TIME)
    v_exp = a_f[:, None] + G_f @ x
    x_exp = x + dt * (a_c[:, None] + G_c @ x)
    C = orc.dinv(P, np.float64) * st.B
    assert np.abs(st.v - v_exp).max() < 1e-10
    assert np.abs(st.x - x_exp).max() < 1e-12
    assert np.abs(C - G_f.reshape(-1)[:, None]).max() < 1e-9
    # F_E' = clamp(I + dt G_f) F_E: with |dt G_f| tiny the singular values stay in
    # the elastic band, so no clamp acts and F_E' = I + dt G_f
    assert np.abs(st.Fe - (eye + dt * G_f.reshape(-1))[:, None]).max() < 1e-10
    # the swapped assignment would be far off
    assert np.abs(v_exp - (a_c[:, None] + G_c @ x)).min() > 1e3 * 1e-10
    assert np.abs(dt * ((a_c - a_f)[:, None] + (G_c - G_f) @ x)).max() > 1e3 * 1e-12
