"""Pins the oracle's substep stages (P2G, grid update, G2P) with closed forms and
conservation laws, so a dropped term, a wrong sign or a transposed operand fails.
CPU only."""
import numpy as np
import pytest

from mpm_snow_sim2d_amd.params import BOUNDARY_STICKY, reference_params, snow_params
from mpm_snow_sim2d_amd.scenes import ParticleSet


def _block(dim, lo, hi, h, v, rng=None, rho=400.0):
    axes = [np.arange(lo[a] + h / 2, hi[a], h) for a in range(dim)]
    pts = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, dim)
    if rng is not None:
        pts = pts + rng.uniform(-0.3 * h, 0.3 * h, pts.shape)
    n = pts.shape[0]
    x = np.ascontiguousarray(pts.T, np.float64)
    vv = np.repeat(np.asarray(v, np.float64)[:, None], n, 1)
    vol = h ** dim
    return ParticleSet(x, vv, np.full(n, rho * vol), np.full(n, vol))


# ---------------------------------------------------------------- grid update (grid_update.glsl)
def _one_node(P, ijk, mom, mass=2.0):
    G = P.num_nodes
    d = P.dim
    W1, H1 = P.WIDTH + 1, P.HEIGHT + 1
    nid = ijk[0] + ijk[1] * W1 + (ijk[2] * W1 * H1 if d == 3 else 0)
    gm = np.zeros(G)
    gmom = np.zeros((d, G))
    gm[nid] = mass
    gmom[:, nid] = np.asarray(mom) * mass
    return gm, gmom, nid


@pytest.mark.parametrize("dim", [2, 3])
def test_grid_update_free_node(orc, dim):
    P = reference_params(dim)
    v = [1.5, -2.0, 0.5][:dim]
    gm, gmom, nid = _one_node(P, [100, 50, 50][:dim], v)
    vc, vf = orc.grid_update(P, gm, gmom)
    want = np.array(v, float)
    want[1] += np.float64(np.float32(P.GRAVITY)) * np.float64(np.float32(P.DELTA_TIME))
    assert np.allclose(vc[:, nid], want, rtol=1e-12)
    assert np.array_equal(vc[:, nid], vf[:, nid])
    # every other node untouched (skipped as light)
    vc[:, nid] = 0
    assert not vc.any()


def test_grid_update_light_node_skipped(orc):
    P = reference_params(2)
    gm, gmom, nid = _one_node(P, [100, 50], [1.0, 1.0], mass=5e-7)   # below 1e-6
    vc, vf = orc.grid_update(P, gm, gmom)
    assert not vc.any() and not vf.any()


@pytest.mark.parametrize("vy", [0.0, 3.0, -0.2])
def test_grid_update_right_wall_projection_and_friction(orc, vy):
    """Node on the right wall (x = W - margin) moving into it: the normal velocity is
    removed (vcol.x = 0) and Coulomb friction takes min(|vt|, mu |dv|) off the
    tangential part."""
    P = reference_params(2)
    P.GRAVITY = 0.0
    vx = 5.0
    gm, gmom, nid = _one_node(P, [P.WIDTH - 2, 50], [vx, vy])
    vc, vf = orc.grid_update(P, gm, gmom)
    assert np.allclose(vc[:, nid], [0.0, vy], atol=1e-12)
    red = min(abs(vy), P.FRICTION * vx)
    want_vfy = vy - np.sign(vy) * red if abs(vy) > 1e-7 else vy
    assert np.allclose(vf[:, nid], [0.0, want_vfy], atol=1e-6)


def test_grid_update_separating_node_untouched(orc):
    """A node already past the wall but moving out of it is not projected."""
    P = reference_params(2)
    P.GRAVITY = 0.0
    gm, gmom, nid = _one_node(P, [P.WIDTH - 1, 50], [-4.0, 1.0])
    vc, vf = orc.grid_update(P, gm, gmom)
    assert np.allclose(vc[:, nid], [-4.0, 1.0]) and np.allclose(vf[:, nid], [-4.0, 1.0])


def test_grid_update_corner_sequential(orc):
    """Bottom-left corner node moving into both walls: bottom is processed before
    left (grid_update.glsl:22-33), both normal components vanish."""
    P = reference_params(2)
    P.GRAVITY = 0.0
    gm, gmom, nid = _one_node(P, [2, 2], [-3.0, -4.0])
    vc, vf = orc.grid_update(P, gm, gmom)
    assert np.allclose(vc[:, nid], [0.0, 0.0], atol=1e-12)
    assert np.allclose(vf[:, nid], [0.0, 0.0], atol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_grid_update_3d_floor_penetration(orc, dim):
    """Node below the floor (y < margin) moving further down: test - min(dist,0) =
    dt*vy, the correction removes exactly vy (grid_update.glsl:48-57)."""
    P = reference_params(dim)
    P.GRAVITY = 0.0
    gm, gmom, nid = _one_node(P, [60, 1, 40][:dim], [0.0, -2.0, 0.0][:dim])
    vc, _ = orc.grid_update(P, gm, gmom)
    assert np.allclose(vc[:, nid], 0.0, atol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_grid_update_sticky(orc, dim):
    P = snow_params(dim, 32, 1e-4)
    P.GRAVITY = 0.0
    for ijk, stuck in (([1, 16, 16], True), ([16, 2, 16], True), ([16, 16, 30], dim == 3),
                       ([16, 29, 16], False), ([16, 16, 16], False)):
        gm, gmom, nid = _one_node(P, ijk[:dim], [0.3, -0.4, 0.2][:dim])
        vc, vf = orc.grid_update(P, gm, gmom)
        if stuck:
            assert not vc[:, nid].any() and not vf[:, nid].any()
        else:
            assert np.allclose(vc[:, nid], [0.3, -0.4, 0.2][:dim])


# ---------------------------------------------------------------- P2G conservation (p2g.glsl)
@pytest.mark.parametrize("dim,order", [(2, 3), (2, 2), (3, 3), (3, 2)])
def test_p2g_conserves_mass_and_momentum(orc, dim, order):
    """Sum_i m_i = Sum_p m_p and Sum_i (mv)_i = Sum_p m_p v_p: the affine (C) and
    stress terms are first moments of w and vanish by linear reproduction."""
    P = reference_params(dim) if order == 3 else snow_params(dim, 32, 1e-4)
    rng = np.random.default_rng(dim * 10 + order)
    if order == 3:
        ps = _block(dim, [40.0] * dim, [48.0] * dim, 0.7, [1.0, -2.0, 0.5][:dim], rng)
    else:
        ps = _block(dim, [0.4] * dim, [0.55] * dim, 0.012, [1.0, -2.0, 0.5][:dim], rng)
    st = orc.OracleState.from_particles(ps, np.float64)
    n = st.n
    st.B = rng.normal(scale=0.5, size=st.B.shape)
    st.Fe = st.Fe + rng.normal(scale=0.02, size=st.Fe.shape)
    st.Fp = st.Fp + rng.normal(scale=0.01, size=st.Fp.shape)
    st.v = rng.normal(size=st.v.shape)
    gm, gmom = orc.p2g(P, st)
    assert np.isclose(gm.sum(), st.mass.sum(), rtol=1e-12)
    assert np.allclose(gmom.sum(1), (st.mass * st.v).sum(1), rtol=1e-9, atol=1e-9 * np.abs(st.mass * st.v).sum())
    assert n > 100


def test_p2g_skips_zero_volume(orc):
    P = reference_params(2)
    ps = _block(2, [40.0, 40.0], [42.0, 42.0], 0.5, [1.0, 0.0])
    st = orc.OracleState.from_particles(ps, np.float64)
    st.vol[:] = 0.0
    gm, gmom = orc.p2g(P, st)
    assert not gm.any() and not gmom.any()


def test_p2g_clips_outside_nodes(orc):
    """A particle next to the grid edge loses the weight of the skipped nodes
    (p2g.glsl:69: skip, not clamp)."""
    P = reference_params(2)
    ps = _block(2, [0.0, 50.0], [0.5, 50.5], 0.5, [0.0, 0.0])
    st = orc.OracleState.from_particles(ps, np.float64)
    gm, _ = orc.p2g(P, st)
    assert gm.sum() < st.mass.sum() * 0.999


# ---------------------------------------------------------------- full substeps
@pytest.mark.parametrize("dim,order", [(2, 3), (2, 2), (3, 2)])
def test_free_fall_closed_form(orc, dim, order):
    """A stress-free block in free fall: every particle follows v_{k+1} = v_k + g dt,
    x_{k+1} = x_k + dt v_{k+1}; C stays 0 and F stays I."""
    if order == 3:
        P = reference_params(dim)
        ps = _block(dim, [60.0] * dim, [66.0] * dim, 0.8, [0.5, 0.0, -0.25][:dim])
    else:
        P = snow_params(dim, 64, 1e-4)
        ps = _block(dim, [0.45] * dim, [0.55] * dim, 0.01, [0.5, 0.0, -0.25][:dim])
    st = orc.OracleState.from_particles(ps, np.float64)
    x0, v0 = st.x.copy(), st.v.copy()
    k = 7
    orc.run(P, st, k)
    g = np.float64(np.float32(P.GRAVITY)) * np.float64(np.float32(P.DELTA_TIME))
    dt = np.float64(np.float32(P.DELTA_TIME))
    v, x = v0.copy(), x0.copy()
    for _ in range(k):
        v[1] += g
        x = x + dt * v
    assert np.allclose(st.v, v, atol=1e-9)
    assert np.allclose(st.x, x, atol=1e-9 * max(1.0, np.abs(x).max()))
    # nodes lighter than mass_eps keep zero velocity (grid_update.glsl:90), which
    # leaves a residue of order (mass_eps / m_p) |v| in B
    assert np.abs(st.B).max() < 1e-7
    eye = np.eye(dim).reshape(-1)[:, None]
    assert np.abs(st.Fe - eye).max() < 1e-9


@pytest.mark.parametrize("order", [2, 3])
def test_affine_field_reproduced(orc, order):
    """With no stress and no gravity an affine velocity field v = v0 + G (x - c) is
    carried through P2G -> grid -> G2P exactly: particle v and C = D^-1 B equal the
    field and its gradient."""
    dim = 2
    P = reference_params(dim) if order == 3 else snow_params(dim, 64, 1e-4)
    P.GRAVITY = 0.0
    P.E = 0.0
    P.derive_lame()
    if order == 3:
        ps = _block(dim, [60.0, 40.0], [70.0, 48.0], 0.5, [0.0, 0.0], np.random.default_rng(0))
        c = np.array([65.0, 44.0])
    else:
        ps = _block(dim, [0.45, 0.45], [0.55, 0.52], 0.006, [0.0, 0.0], np.random.default_rng(0))
        c = np.array([0.5, 0.48])
    Gm = np.array([[0.1, -0.7], [0.7, 0.05]])
    v0 = np.array([0.3, -0.2])
    st = orc.OracleState.from_particles(ps, np.float64)
    st.v = v0[:, None] + Gm @ (st.x - c[:, None])
    st.B = np.repeat((Gm / orc.dinv(P, np.float64)).reshape(-1)[:, None], st.n, 1)
    x_old = st.x.copy()
    orc.run(P, st, 1)
    vel_field = v0[:, None] + Gm @ (x_old - c[:, None])
    assert np.allclose(st.v, vel_field, atol=1e-10)
    C = orc.dinv(P, np.float64) * st.B
    assert np.allclose(C, Gm.reshape(-1)[:, None], atol=1e-9)
    dt = np.float64(np.float32(P.DELTA_TIME))
    assert np.allclose(st.x, x_old + dt * vel_field, atol=1e-10)


@pytest.mark.parametrize("dim", [2, 3])
def test_block_at_rest_stays_at_rest(orc, dim):
    P = snow_params(dim, 48, 1e-4)
    P.GRAVITY = 0.0
    ps = _block(dim, [0.4] * dim, [0.5] * dim, 0.008, [0.0] * dim)
    st = orc.OracleState.from_particles(ps, np.float64)
    x0 = st.x.copy()
    orc.run(P, st, 5)
    assert np.abs(st.v).max() < 1e-12
    assert np.abs(st.x - x0).max() < 1e-12


def test_plastic_split_preserves_total_gradient(orc):
    """G2P splits F_hat = (I + dt C) F_E into F_E' F_P' with F_E' F_P' = F_hat F_P
    (g2p.glsl:74-93), while F_E' singular values lie in [1-theta_c, 1+theta_s]."""
    P = reference_params(2)
    P.GRAVITY = 0.0
    ps = _block(2, [80.0, 40.0], [90.0, 46.0], 0.6, [0.0, 0.0], np.random.default_rng(4))
    st = orc.OracleState.from_particles(ps, np.float64)
    st.v = np.stack([-3.0 * (st.x[0] - 85.0), 2.0 * (st.x[1] - 43.0)])     # strong compression / stretch
    st.B = np.repeat((np.diag([-3.0, 2.0]) / 3.0).reshape(-1)[:, None], st.n, 1)
    Fe0, Fp0 = st.Fe.copy(), st.Fp.copy()
    orc.run(P, st, 1)
    C = 3.0 * st.B
    dt = np.float64(np.float32(P.DELTA_TIME))
    for p in range(0, st.n, 7):
        M = np.eye(2) + dt * C[:, p].reshape(2, 2)
        Fh = M @ Fe0[:, p].reshape(2, 2)
        lhs = st.Fe[:, p].reshape(2, 2) @ st.Fp[:, p].reshape(2, 2)
        assert np.allclose(lhs, Fh @ Fp0[:, p].reshape(2, 2), atol=1e-10)
        s = np.linalg.svd(st.Fe[:, p].reshape(2, 2))[1]
        assert s.min() >= 1 - np.float32(P.THETA_C) - 1e-9 and s.max() <= 1 + np.float32(P.THETA_S) + 1e-9


@pytest.mark.parametrize("dim", [2, 3])
def test_f32_oracle_tracks_f64(orc, dim):
    """The FP32 restatement stays within 1e-4 (relative to the scale) of its FP64 twin
    over 20 substeps of a colliding pair of blocks."""
    P = snow_params(dim, 32, 1e-4)
    rng = np.random.default_rng(5)
    a = _block(dim, [0.3] * dim, [0.42] * dim, 0.012, [2.0, 0.0, 0.0][:dim], rng)
    b = _block(dim, [0.44] + [0.32] * (dim - 1), [0.56] + [0.44] * (dim - 1), 0.012, [-2.0, 0.0, 0.0][:dim], rng)
    ps = a.concat(b)
    s64 = orc.OracleState.from_particles(ps, np.float64)
    s32 = orc.OracleState.from_particles(ps, np.float32)
    orc.run(P, s64, 20)
    orc.run(P, s32, 20)
    assert np.abs(s32.x - s64.x).max() < 1e-5
    assert np.abs(s32.v - s64.v).max() < 1e-3 * np.abs(s64.v).max()
    assert np.abs(s32.Fe - s64.Fe).max() < 1e-4


@pytest.mark.parametrize("dim,scheme", [(2, "mls"), (3, "mls"), (2, "apic")])
def test_multithreaded_baseline_matches_serial_oracle(orc, dim, scheme):
    """The all-cores CPU baseline (OpenMP build of the same body) runs the same
    substep, of the GLSL path's MLS scheme and of the CPU path's APIC scheme with
    the explicit grid force: after 10 substeps it agrees with the serial FP32
    oracle up to the order of the grid's float sums."""
    from mpm_snow_sim2d_amd.params import SCHEME_APIC_FORCE
    P = snow_params(dim, 32, 1e-4)
    if scheme == "apic":
        P.scheme = SCHEME_APIC_FORCE
    rng = np.random.default_rng(6)
    a = _block(dim, [0.3] * dim, [0.42] * dim, 0.012, [2.0, 0.0, 0.0][:dim], rng)
    b = _block(dim, [0.44] + [0.32] * (dim - 1), [0.56] + [0.44] * (dim - 1), 0.012, [-2.0, 0.0, 0.0][:dim], rng)
    ps = a.concat(b)
    ser = orc.OracleState.from_particles(ps, np.float32)
    par = orc.OracleState.from_particles(ps, np.float32)
    orc.run(P, ser, 10)
    orc.run_omp(P, par, 10)
    assert orc.omp_threads() >= 1
    assert np.abs(par.x - ser.x).max() < 1e-5
    assert np.abs(par.v - ser.v).max() < 1e-3 * np.abs(ser.v).max()
    assert np.abs(par.Fe - ser.Fe).max() < 1e-4
