"""The learned surrogate on the device, through the C-ABI.

* Network: forward output and parameter gradients of the BF16 matrix-core U-Net
  against the FP64 oracle (oracle/unet_ref.py) fed the same BF16-rounded weights
  and inputs.  Bounds, relative L2 over each tensor: output 1e-2, loss 1e-2, every
  parameter block's gradient 2e-2 with cosine similarity >= 0.9995 (BF16 keeps 8
  significant bits; activations and activation gradients are rounded to it after
  every layer); the head bias, on which the projected loss does not depend, has a
  gradient of zero up to rounding.
* Capture: node features against a NumPy P2G of the same particles (FP32 atomics
  against FP64 sums: relative 1e-5); the velocity and displacement targets, as
  transfers and after 8 steps of their least-squares refinement, against a NumPy
  restatement (relative 2e-4 of the field's largest value), which also shows the
  refinement lowers the particles' velocity fit residual and keeps the fields'
  mass-weighted means.
* Macro step: ballistic advance of a stress-free body in free fall against its
  closed form (trapezoid positions, or the solver's own free-fall recurrence when
  the displacement moves the particles), fallback to plain substeps at the walls.
* Training: Adam fits a fixed batch with one image per sample and with the
  velocity + displacement pair.
* Ablation: a network trained on solver trajectories predicts the velocity of
  held-out colliding scenes better than the method without it (the ballistic grid
  velocity, i.e. a zero network output), and its hybrid rollout lands closer to
  the plain solver.
"""
import numpy as np
import pytest

from mpm_snow_sim2d_amd.params import Options, snow_params
from mpm_snow_sim2d_amd.scenes import ParticleSet, colliding_scene, jitter_lattice, _set
from oracle import unet_ref as U

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _net_problem(rng, sur, nb, h, w):
    params = sur.get_params().astype(np.float64)
    p = U.unpack(params.copy())
    for name in U.NAMES:      # the device convolves with BF16 copies of the weights
        p[name + ".w"] = U.bf16_round(p[name + ".w"])
    x = rng.normal(size=(nb, h, w, 32)).astype(np.float32)
    x[..., 8:] = 0.0
    bits = U.to_bf16_bits(x).reshape(nb, h, w, 32)
    xr = U.from_bf16_bits(bits).astype(np.float64)
    m = rng.uniform(0.0, 1.0, size=(nb, h * w)) * (rng.uniform(size=(nb, h * w)) > 0.4)
    wm = (m / m.sum(1, keepdims=True)).astype(np.float32)
    target = rng.normal(scale=0.5, size=(nb, h * w, 2)).astype(np.float32)
    return p, bits, xr, wm, target


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def test_network_forward_matches_oracle():
    from mpm_snow_sim2d_amd.surrogate import Surrogate
    rng = np.random.default_rng(0)
    nb, h, w = 2, 32, 32
    _flush("[surrogate] forward 2x32x32")
    with Surrogate(h, w, max_batch=nb, seed=1) as sur:
        # a head large enough that the output is not dominated by rounding
        prm = sur.get_params()
        o, _ = U.layout()[0]["head.w"]
        prm[o:o + 64] = rng.normal(scale=0.5, size=64)
        sur.set_params(prm)
        p, bits, xr, wm, _ = _net_problem(rng, sur, nb, h, w)
        out = sur.forward(bits, wm)
    ref = U.forward(p, xr, wm.astype(np.float64))
    err = _rel(out, ref)
    _flush(f"    rel L2 {err:.3e}")
    assert err < 1e-2


@pytest.mark.parametrize("kernel", ["stream", "lds"])
def test_network_forward_batch_matches_oracle(kernel):
    """16 images of 32x32: with the LDS-tiled convolutions (4-row tiles) and with
    the register-streaming ones, against the FP64 oracle."""
    from mpm_snow_sim2d_amd.surrogate import CONV_LDS, CONV_STREAM, Surrogate
    rng = np.random.default_rng(5)
    nb, h, w = 16, 32, 32
    _flush(f"[surrogate] forward {nb}x{h}x{w} {kernel}")
    with Surrogate(h, w, max_batch=nb, seed=4) as sur:
        sur.set_conv(CONV_LDS if kernel == "lds" else CONV_STREAM)
        prm = sur.get_params()
        o, _ = U.layout()[0]["head.w"]
        prm[o:o + 64] = rng.normal(scale=0.5, size=64)
        sur.set_params(prm)
        p, bits, xr, wm, _ = _net_problem(rng, sur, nb, h, w)
        out = sur.forward(bits, wm)
    ref = U.forward(p, xr, wm.astype(np.float64))
    err = _rel(out, ref)
    _flush(f"    rel L2 {err:.3e}")
    assert err < 1e-2


@pytest.mark.parametrize("nb,size", [(1, 256), (8, 128), (8, 256)])
def test_lds_convolutions_match_streaming(nb, size):
    """The LDS-tiled forward convolutions at the shapes that select each tile height
    (4 rows: one 256^2 image; 8 rows: 8 images of 128^2; 16 rows: 8 images of
    256^2) against the register-streaming kernels (checked against the oracle at
    small sizes): equal up to FP32 accumulation order and the BF16 roundings of
    the activations that order can flip."""
    from mpm_snow_sim2d_amd.surrogate import CONV_AUTO, CONV_STREAM, Surrogate
    rng = np.random.default_rng(6)
    _flush(f"[surrogate] lds vs streaming {nb}x{size}x{size}")
    with Surrogate(size, size, max_batch=nb, seed=5) as sur:
        prm = sur.get_params()
        o, _ = U.layout()[0]["head.w"]
        prm[o:o + 64] = rng.normal(scale=0.5, size=64)
        sur.set_params(prm)
        x = rng.normal(size=(nb, size, size, 32)).astype(np.float32)
        x[..., 8:] = 0.0
        bits = U.to_bf16_bits(x).reshape(nb, size, size, 32)
        m = rng.uniform(0.0, 1.0, size=(nb, size * size))
        wm = (m / m.sum(1, keepdims=True)).astype(np.float32)
        sur.set_conv(CONV_STREAM)
        ref = sur.forward(bits, wm)
        sur.set_conv(CONV_AUTO)
        outs = [sur.forward(bits, wm) for _ in range(2)]
    assert np.array_equal(outs[0], outs[1])   # no atomics: the LDS path is deterministic
    err = _rel(outs[0], ref)
    _flush(f"    rel L2 {err:.3e}")
    assert err < 5e-3


def test_network_gradient_matches_oracle():
    from mpm_snow_sim2d_amd.surrogate import Surrogate
    rng = np.random.default_rng(1)
    nb, h, w = 2, 32, 32
    _flush("[surrogate] gradient 2x32x32")
    with Surrogate(h, w, max_batch=nb, seed=2) as sur:
        prm = sur.get_params()
        o, _ = U.layout()[0]["head.w"]
        prm[o:o + 64] = rng.normal(scale=0.5, size=64)
        sur.set_params(prm)
        p, bits, xr, wm, target = _net_problem(rng, sur, nb, h, w)
        loss, g = sur.gradient(bits, wm, target)
    lref, gref = U.loss_and_grad(p, xr, wm.astype(np.float64), target.astype(np.float64))
    _flush(f"    loss {loss:.6f} oracle {lref:.6f}")
    assert abs(loss - lref) <= 1e-2 * abs(lref)
    gd = U.unpack(g.astype(np.float64))
    # the projection removes any constant from the output, so the loss does not
    # depend on the head bias: its gradient vanishes identically
    assert np.abs(gd["head.b"]).max() <= 1e-5 * np.linalg.norm(gd["head.w"])
    for k in gref:
        if k == "head.b":
            continue
        a, b = gd[k].reshape(-1), gref[k].reshape(-1)
        cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
        err = _rel(a, b)
        _flush(f"    {k:8s} rel {err:.3e} cos {cos:.5f}")
        assert err < 2e-2 and cos >= 0.9995, k


@pytest.mark.parametrize("flags,refine", [(0, 0), (3, 8)])
def test_training_steps_reduce_loss(flags, refine):
    """Adam on the device fits a fixed batch of captured samples (velocity only, and
    velocity + displacement with refined targets)."""
    from mpm_snow_sim2d_amd.surrogate import Surrogate, generate
    P = snow_params(2, 32, 4e-4)
    scenes = [(P, colliding_scene(50 + i, 1500)) for i in range(2)]
    _flush(f"[surrogate] fixed-batch training flags={flags} refine={refine}")
    with Surrogate(32, 32, max_batch=4, capacity=24, seed=3) as sur:
        sur.set_macro(flags)
        sur.set_targets(refine)
        # the last samples of 96 substeps, after the bodies have met
        slots = [ss[-2:] for ss in generate(sur, scenes, 96, 8)]
        slots = [s for ss in slots for s in ss]
        l0 = sur.loss(slots)
        assert l0 > 1e-3
        for _ in range(60):
            sur.train(slots, 2e-3)
        l1 = sur.loss(slots)
    _flush(f"[surrogate] fixed-batch loss {l0:.4f} -> {l1:.4f}")
    assert np.isfinite(l1) and l1 < 0.5 * l0


def _p2g_numpy(P, x, vals, m, h, w):
    """sum_p w_ip m_p vals_p over nodes [0,w) x [0,h) (quadratic B-spline, FP64)."""
    inv = 1.0 / P.dx
    acc = np.zeros((vals.shape[0] + 1, h * w))
    xg = x.astype(np.float64) * inv
    base = np.floor(xg - 0.5).astype(np.int64)
    for a in range(3):
        for b in range(3):
            i, j = base[0] + a, base[1] + b
            di, dj = np.abs(xg[0] - i), np.abs(xg[1] - j)
            wx = np.where(di < 0.5, 0.75 - di * di, np.where(di < 1.5, 0.5 * (1.5 - di) ** 2, 0.0))
            wy = np.where(dj < 0.5, 0.75 - dj * dj, np.where(dj < 1.5, 0.5 * (1.5 - dj) ** 2, 0.0))
            wgt = wx * wy * m
            ok = (i >= 0) & (i < w) & (j >= 0) & (j < h)
            q = (j * w + i)[ok]
            np.add.at(acc[0], q, wgt[ok])
            for c in range(vals.shape[0]):
                np.add.at(acc[c + 1], q, (wgt * vals[c])[ok])
    return acc


def _gather_numpy(P, x, field, h, w):
    """sum_i w_ip field_i over nodes [0,w) x [0,h) (quadratic B-spline, FP64)."""
    inv = 1.0 / P.dx
    xg = x.astype(np.float64) * inv
    base = np.floor(xg - 0.5).astype(np.int64)
    out = np.zeros((field.shape[0], x.shape[1]))
    for a in range(3):
        for b in range(3):
            i, j = base[0] + a, base[1] + b
            di, dj = np.abs(xg[0] - i), np.abs(xg[1] - j)
            wx = np.where(di < 0.5, 0.75 - di * di, np.where(di < 1.5, 0.5 * (1.5 - di) ** 2, 0.0))
            wy = np.where(dj < 0.5, 0.75 - dj * dj, np.where(dj < 1.5, 0.5 * (1.5 - dj) ** 2, 0.0))
            ok = (i >= 0) & (i < w) & (j >= 0) & (j < h)
            q = np.where(ok, j * w + i, 0)
            out += np.where(ok, wx * wy, 0.0) * field[:, q]
    return out


def _fit_numpy(P, x, y, m, h, w, iters=8):
    """Mass-weighted least-squares grid fields of the particle values y (c, n): the
    P2G transfer, then `iters` steps t += P2G(m (y - G2P(t))) / m_i."""
    acc = _p2g_numpy(P, x, y, m, h, w)
    mass = acc[0]
    inv = np.where(mass > 0, 1.0 / np.where(mass > 0, mass, 1.0), 0.0)
    t = acc[1:] * inv
    for _ in range(iters):
        r = y - _gather_numpy(P, x, t, h, w)
        t = t + _p2g_numpy(P, x, r, m, h, w)[1:] * inv
    return t


@pytest.mark.parametrize("refine", [0, 8])
def test_capture_matches_numpy(refine):
    from mpm_snow_sim2d_amd.solver import GPUSolver
    from mpm_snow_sim2d_amd.surrogate import Surrogate
    P = snow_params(2, 64, 2e-4)
    ps = colliding_scene(5, 3000)
    k = 8
    _flush(f"[surrogate] capture n={ps.n} refine={refine}")
    with Surrogate(64, 64, max_batch=1, capacity=2) as sur:
        sur.set_targets(refine)
        s = GPUSolver(P, ps, Options(sort_every=4, tile_particles=256))
        s.Step(6)
        st0 = s.get_state()
        sur.record(s, 0, 1, 0, k)
        s.Step(k)
        st1 = s.get_state()
        sur.record(s, 0, 1, 1, k)
        data, sc = sur.get_sample(1)
        s.close()
    m = ps.mass.astype(np.float64)
    F = st0["F"].astype(np.float64)
    ls = 0.5 * np.log(np.abs(F[0] * F[3] - F[1] * F[2]))
    acc = _p2g_numpy(P, st0["x"], np.stack([st0["v"][0], st0["v"][1], st0["Jp"], ls]), m, 64, 64)
    y = np.concatenate([st1["v"].astype(np.float64), st1["x"].astype(np.float64) - st0["x"]])
    tgt = _fit_numpy(P, st0["x"], y, m, 64, 64, refine)
    mass = acc[0]
    occ = mass > 1e-6 * mass.max()
    assert np.allclose(data[0], mass, rtol=1e-5, atol=1e-6 * mass.max())
    for c in range(4):
        ref = acc[c + 1][occ] / mass[occ]
        # log|det F_E| / 2 of a nearly undeformed body is ~1e-4: FP32 rounding of
        # det F_E (6e-8) bounds it absolutely
        tol = max(1e-4 * np.abs(ref).max(), 1e-6)
        assert np.abs(data[c + 1][occ] - ref).max() <= tol, c
    for c in range(4):
        ref = tgt[c][occ]
        assert np.abs(data[5 + c][occ] - ref).max() <= 2e-4 * np.abs(ref).max(), c
    # the refinement fits the particles' velocities better than the plain transfer
    # (the displacements of these still nearly rigid bodies sit at the rounding
    # floor either way) and keeps each field's mass-weighted sum (all stencils lie
    # inside the image here)
    pic = _p2g_numpy(P, st0["x"], y, m, 64, 64)[1:] / np.where(mass > 0, mass, 1.0)
    res_fit = np.sum(m * (y[:2] - _gather_numpy(P, st0["x"], tgt[:2], 64, 64)) ** 2)
    res_pic = np.sum(m * (y[:2] - _gather_numpy(P, st0["x"], pic[:2], 64, 64)) ** 2)
    if refine:
        assert res_fit < 0.5 * res_pic
    assert np.allclose((mass * tgt).sum(1), (m * y).sum(1), rtol=1e-9, atol=1e-12)
    assert abs(sc["particle_mass"] - m.sum()) <= 1e-5 * m.sum()
    assert abs(sc["T"] - k * np.float32(P.DELTA_TIME)) < 1e-9
    assert sc["band_mass"] == 0.0


@pytest.mark.parametrize("flags", [0, 3])
def test_ballistic_macro_step_free_fall(flags):
    """A stress-free body with uniform velocity: v' = v + g T; x' = x + T v + g T^2 / 2
    (trapezoid), or, moved by the displacement, the solver's own free-fall
    recurrence x' = x + T v + g dt^2 k (k + 1) / 2; F and J_P unchanged."""
    from mpm_snow_sim2d_amd.solver import GPUSolver
    from mpm_snow_sim2d_amd.surrogate import ADVANCE_BALLISTIC, Surrogate
    P = snow_params(2, 64, 2e-4)
    rng = np.random.default_rng(4)
    c = np.array([0.5, 0.55])
    pts = jitter_lattice(c - 0.08, c + 0.08, 0.004, rng, lambda p: ((p - c) ** 2).sum(1) <= 0.08 ** 2)
    ps = _set(pts, (0.7, -0.3), 0.004 ** 2, 400 * 0.004 ** 2)
    k = 8
    with Surrogate(64, 64, max_batch=1, capacity=0) as sur:
        sur.set_macro(flags)
        s = GPUSolver(P, ps, Options(sort_every=4))
        st0 = s.get_state()
        used, res = sur.advance(s, k, ADVANCE_BALLISTIC, 1e-3)
        st1 = s.get_state()
        s.close()
    assert used == ADVANCE_BALLISTIC and res <= 1e-3
    T = k * np.float32(P.DELTA_TIME)
    g = np.array([0.0, P.GRAVITY])
    v0 = st0["v"].astype(np.float64)
    assert np.abs(st1["v"] - (v0 + g[:, None] * T)).max() < 1e-5
    dt = np.float32(P.DELTA_TIME)
    fall = dt * dt * (k * (k + 1) / 2) if flags else 0.5 * T * T
    xr = st0["x"] + T * v0 + g[:, None] * fall
    assert np.abs(st1["x"] - xr).max() < 1e-6
    assert np.abs(st1["F"] - st0["F"]).max() < 1e-5
    assert np.abs(st1["Jp"] - st0["Jp"]).max() < 1e-6


@pytest.mark.parametrize("flags", [0, 3])
def test_macro_step_matches_restatement(flags):
    """The ballistic macro step of a deforming scene (two bodies meeting: non-uniform
    velocity, C and F_E) against the NumPy restatement of k_macro_g2p
    (tests/surrogate_floor.py) fed the transfer of the same state in FP64: x, v, C,
    F_E and J_P, both forms (trapezoid / displacement)."""
    import surrogate_floor as SF
    from mpm_snow_sim2d_amd.solver import GPUSolver
    from mpm_snow_sim2d_amd.surrogate import ADVANCE_BALLISTIC, Surrogate
    from oracle import oracle as orc
    P = snow_params(2, 64, 2e-4)
    ps = colliding_scene(5, 3000)
    k = 8
    _flush(f"[surrogate] macro step vs restatement flags={flags}")
    with Surrogate(64, 64, max_batch=1, capacity=0) as sur:
        sur.set_macro(flags)
        s = GPUSolver(P, ps, Options(sort_every=4, tile_particles=256))
        s.Step(128)                 # the bodies have met and deform
        st0 = s.get_state()
        used, res = sur.advance(s, k, ADVANCE_BALLISTIC, 1e-3)
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
    img = SF.Image(P, 64, 64)
    sten = img.stencil(ost.x)
    m = ost.mass
    dt = float(np.float32(P.DELTA_TIME))
    T, g = k * dt, float(np.float32(P.GRAVITY))
    vf, _, _ = img.transfer(sten, m, ost.v)
    vf[1] += g * T
    uf = None
    if flags:
        uf, _, _ = img.transfer(sten, m, T * ost.v)
        uf[1] += g * dt * dt * k * (k + 1) / 2
    SF.macro_step(P, img, ost, k, vf, uf)
    vscale = np.abs(ost.v).max()
    assert np.abs(st1["x"] - ost.x).max() < 1e-6
    assert np.abs(st1["v"] - ost.v).max() < 1e-4 * vscale
    C = dinv * ost.B
    assert np.abs(st1["C"] - C).max() < 1e-3 * max(np.abs(C).max(), 1.0)
    assert np.abs(st1["F"] - ost.Fe).max() < 1e-5
    Jp = ost.Fp[0] * ost.Fp[3] - ost.Fp[1] * ost.Fp[2]
    assert np.abs(st1["Jp"] - Jp).max() < 1e-5
    # the step is not trivial: the bodies deform
    assert np.abs(st1["F"] - st0["F"]).max() > 1e-4


def test_option_setters_reject_bad_values():
    """mpm_surrogate_set_macro / set_targets refuse values outside their range and
    leave the surrogate usable."""
    from mpm_snow_sim2d_amd.surrogate import MACRO_F, MACRO_X, Surrogate
    with Surrogate(32, 32, max_batch=1, capacity=1) as sur:
        for bad in (4, -1, 8):
            with pytest.raises(RuntimeError):
                sur.set_macro(bad)
        for bad in (-1, 65):
            with pytest.raises(RuntimeError):
                sur.set_targets(bad)
        sur.set_macro(MACRO_X | MACRO_F)
        sur.set_targets(64)
        sur.set_macro(0)
        sur.set_targets(0)


def test_advance_falls_back_at_walls():
    from mpm_snow_sim2d_amd.solver import GPUSolver
    from mpm_snow_sim2d_amd.surrogate import ADVANCE_NETWORK, ADVANCE_PLAIN, Surrogate
    P = snow_params(2, 64, 2e-4)
    rng = np.random.default_rng(5)
    c = np.array([0.5, 0.1])
    pts = jitter_lattice(c - 0.06, c + 0.06, 0.004, rng, lambda p: ((p - c) ** 2).sum(1) <= 0.06 ** 2)
    ps = _set(pts, (0.0, -1.0), 0.004 ** 2, 400 * 0.004 ** 2)
    k = 8
    with Surrogate(64, 64, max_batch=1, capacity=0) as sur:
        a = GPUSolver(P, ps, Options(sort_every=4))
        b = GPUSolver(P, ps, Options(sort_every=4))
        used, res = sur.advance(a, k, ADVANCE_NETWORK, 1e-3)
        b.Step(k)
        sa, sb = a.get_state(), b.get_state()
        a.close()
        b.close()
    assert used == ADVANCE_PLAIN and res > 1e-3
    assert np.abs(sa["x"] - sb["x"]).max() < 1e-6


def test_plain_substeps_after_macro_step_are_refused():
    """After a macro step the solver refuses plain substeps (mpm_step, a PLAIN
    advance, a failed residual test), binning, grid resets and state uploads with
    MPM_E_STATE, and it runs no kernel for them; its state stays readable, and a
    solver created from that state steps like one created from the same arrays."""
    from mpm_snow_sim2d_amd._lib import MPM_E_STATE, MPMError
    from mpm_snow_sim2d_amd.solver import GPUSolver
    from mpm_snow_sim2d_amd.surrogate import ADVANCE_BALLISTIC, ADVANCE_PLAIN, Surrogate
    P = snow_params(2, 64, 2e-4)
    rng = np.random.default_rng(6)
    c = np.array([0.5, 0.55])
    pts = jitter_lattice(c - 0.08, c + 0.08, 0.004, rng, lambda p: ((p - c) ** 2).sum(1) <= 0.08 ** 2)
    ps = _set(pts, (0.5, -0.2), 0.004 ** 2, 400 * 0.004 ** 2)
    k = 8
    with Surrogate(64, 64, max_batch=1, capacity=0) as sur:
        s = GPUSolver(P, ps, Options(sort_every=4))
        used, _ = sur.advance(s, k, ADVANCE_BALLISTIC, 1.0)
        assert used == ADVANCE_BALLISTIC
        st = s.get_state()
        refused = [lambda: s.Step(1), lambda: sur.advance(s, k, ADVANCE_PLAIN, 1.0),
                   lambda: sur.advance(s, k, ADVANCE_BALLISTIC, -1.0),     # residual test fails
                   s.sort, s.reset, s.p2g, lambda: s.set_state(**st)]
        for call in refused:
            with pytest.raises(MPMError) as err:
                call()
            assert err.value.code == MPM_E_STATE
        st2 = s.get_state()
        for key in st:
            assert np.array_equal(st[key], st2[key]), key
        # a further macro step is still taken
        used, _ = sur.advance(s, k, ADVANCE_BALLISTIC, 1.0)
        assert used == ADVANCE_BALLISTIC
        s.close()
    # the way on: a new solver from the read-out state
    moved = ParticleSet(st["x"], st["v"], ps.mass, ps.vol)
    outs = []
    for _ in range(2):
        with GPUSolver(P, moved, Options(sort_every=4), C=st["C"], F=st["F"], Jp=st["Jp"]) as t:
            t.Step(k)
            outs.append(t.get_state())
            assert t.check_finite() == 0
    # the two runs differ only in the order of float atomics at shared tile-halo
    # nodes: natural units (C against max|v| / dx), the batch tests' bounds
    from tests.helpers import state_errors
    err = state_errors(outs[0], outs[1], float(P.dx))
    for key, (mx, med) in err.items():
        assert mx < 1e-3 and med < 1e-5, (key, mx, med)


def test_trained_network_beats_ballistic():
    """Ablation: train on trajectories of colliding scenes (64^2 grid), then on
    held-out scenes compare the network with the zero-output (ballistic) method:
    lower velocity loss, and a hybrid rollout whose positions land closer to the
    plain solver's."""
    from mpm_snow_sim2d_amd.surrogate import (ADVANCE_BALLISTIC, ADVANCE_NETWORK, ADVANCE_PLAIN, Surrogate, fit,
                                              generate, rollout)
    cells, dt, k = 64, 2e-4, 8
    P = snow_params(2, cells, dt)
    train_scenes = [(P, colliding_scene(100 + i, 4000)) for i in range(48)]
    test_scenes = [(P, colliding_scene(900 + i, 4000)) for i in range(6)]
    nsub = 240
    cap = (len(train_scenes) + len(test_scenes)) * (nsub // k)
    with Surrogate(cells, cells, max_batch=16, capacity=cap, seed=7, d_ref=0.1) as sur:
        tr = generate(sur, train_scenes, nsub, k, wall_tol=1e-3)
        te = generate(sur, test_scenes, nsub, k, first_slot=len(train_scenes) * (nsub // k), wall_tol=1e-3)
        tr_slots = [s for ss in tr for s in ss]
        te_slots = [s for ss in te for s in ss]
        _flush(f"    samples: train {len(tr_slots)} held-out {len(te_slots)}")
        base = np.mean([sur.baseline_loss(te_slots[i:i + 16]) for i in range(0, len(te_slots), 16)])
        before = np.mean([sur.loss(te_slots[i:i + 16]) for i in range(0, len(te_slots), 16)])
        fit(sur, tr_slots, steps=800, batch=16, lr=1e-3, log_every=100, log=_flush)
        after = np.mean([sur.loss(te_slots[i:i + 16]) for i in range(0, len(te_slots), 16)])
        _flush(f"    held-out loss: ballistic {base:.4f}  untrained {before:.4f}  trained {after:.4f}")
        assert after < 0.6 * base

        errs = {ADVANCE_NETWORK: [], ADVANCE_BALLISTIC: []}
        for Pq, ps in test_scenes:
            ref = rollout(None, Pq, ps, nsub, k, ADVANCE_PLAIN)["x"]
            for mode in errs:
                r = rollout(sur, Pq, ps, nsub, k, mode, tol=1e-3)
                _flush(f"    mode {mode}: macro steps by mode {r['counts']}")
                assert r["nonfinite"] == 0
                errs[mode].append(float(np.linalg.norm(r["x"] - ref, axis=1).mean()))
        net, bal = np.mean(errs[ADVANCE_NETWORK]), np.mean(errs[ADVANCE_BALLISTIC])
        _flush(f"    mean position error after {nsub} substeps: network {net:.3e} m  ballistic {bal:.3e} m")
        assert net < bal
