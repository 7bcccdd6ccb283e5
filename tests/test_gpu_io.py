"""Snapshots and frames on the device.

* A snapshot holds exactly the solver's state (mpm_get_state) and substep count,
  and a solver loaded from it continues like the original (20 more substeps,
  within the float-atomic order bound of the parity tests).
* Frames: the white pixels are exactly the particles projected with the
  reference's cameras (glOrtho(0, W, 0, H) in 2D; glm::perspective(45 deg) and
  glm::lookAt from the orbiting eye in 3D, src/main.cpp:56-79), restated here in
  NumPy; the border / wireframe box is grey.
"""
import math

import numpy as np
import pytest

from mpm_snow_sim2d_amd.params import Options, reference_params, snow_params
from mpm_snow_sim2d_amd.scenes import reference_snowballs
from tests.helpers import make_case, state_errors

pytestmark = pytest.mark.gpu


def test_snapshot_save_load_continue(tmp_path):
    from mpm_snow_sim2d_amd.snapshot import read_snapshot
    from mpm_snow_sim2d_amd.solver import GPUSolver
    P, ps = make_case("snow3d_quad_sticky")
    a = GPUSolver(P, ps, Options(sort_every=8))
    a.Step(10)
    path = str(tmp_path / "a.snap")
    a.save_snapshot(path)
    st = a.get_state()
    P2, snap, steps = read_snapshot(path)
    assert steps == 10 and bytes(P2) == bytes(P)
    for k in ("x", "v", "C", "F", "Jp"):
        assert np.array_equal(snap[k].reshape(st[k].shape), st[k]), k
    assert np.array_equal(snap["mass"], ps.mass) and np.array_equal(snap["vol"], ps.vol)
    b = GPUSolver.from_snapshot(path, Options(sort_every=8))
    a.Step(20)
    b.Step(20)
    err = state_errors(b.get_state(), a.get_state(), P.dx)
    for k in err:
        assert err[k][0] < 1e-3 and err[k][1] < 1e-5, (k, err[k])
    b.save_snapshot(path)
    assert read_snapshot(path)[2] == 30
    a.close()
    b.close()


def _ortho_pixels(x, P, w, h):
    X = np.floor((x[0].astype(np.float32) / np.float32(P.WIDTH * P.dx) * np.float32(2) - 1) * np.float32(0.5) * w
                 + np.float32(0.5) * w)
    Y = np.floor((x[1].astype(np.float32) / np.float32(P.HEIGHT * P.dx) * np.float32(2) - 1) * np.float32(0.5) * h
                 + np.float32(0.5) * h)
    return X.astype(int), Y.astype(int)


def _white(img):
    return (img == 255).all(-1)


def test_frame_2d_matches_ortho_projection(tmp_path):
    from mpm_snow_sim2d_amd.solver import GPUSolver, write_png
    P = reference_params(2)
    ps = reference_snowballs(2, seed=0, num_particles=3000)
    s = GPUSolver(P, ps)
    s.Step(34)
    x = s.get_state()["x"]
    w, h = 1200, 600
    img = s.render_frame(w, h)
    X, Y = _ortho_pixels(x, P, w, h)
    expect = np.zeros((h, w), bool)
    ok = (X >= 0) & (X < w) & (Y >= 0) & (Y < h)
    expect[h - 1 - Y[ok], X[ok]] = True
    got = _white(img)
    # a rounding step of the projection may move a point by one pixel: compare with
    # a one-pixel tolerance in both directions
    assert got.sum() > 0.9 * len(np.unique(np.stack([X, Y]), axis=1).T)
    pad = np.zeros((h + 2, w + 2), bool)
    pad[1:-1, 1:-1] = expect
    near = np.zeros_like(expect)
    for dy in (0, 1, 2):
        for dx in (0, 1, 2):
            near |= pad[dy:dy + h, dx:dx + w]
    assert not (got & ~near).any()
    # grey border at BORDER_MARGIN cells from the left edge
    bx = int(math.floor(P.BORDER_MARGIN / P.WIDTH * w))
    col = img[h // 2, bx - 1:bx + 3]
    assert (col == 128).all(-1).any()
    write_png(str(tmp_path / "f.png"), img)
    s.close()


def _perspective_pixels(x, P, cam, w, h):
    Wd, Hd, Dd = P.WIDTH * P.dx, P.HEIGHT * P.dx, P.DEPTH * P.dx
    ctr = np.array([Wd / 2, Hd / 2, Dd / 2])
    rad, el = math.radians(cam["angle_deg"]), math.radians(cam["elevation_deg"])
    eye = ctr + cam["distance"] * np.array([math.cos(el) * math.sin(rad), math.sin(el), math.cos(el) * math.cos(rad)])
    f = (ctr - eye) / np.linalg.norm(ctr - eye)
    sd = np.cross(f, [0, 1, 0])
    sd /= np.linalg.norm(sd)
    u = np.cross(sd, f)
    view = np.eye(4)
    view[0, :3], view[1, :3], view[2, :3] = sd, u, -f
    view[:3, 3] = [-sd @ eye, -u @ eye, f @ eye]
    t = 1 / math.tan(math.radians(45) / 2)
    zn, zf = 1.0, 1000.0
    proj = np.zeros((4, 4))
    proj[0, 0], proj[1, 1] = t / (w / h), t
    proj[2, 2], proj[2, 3], proj[3, 2] = -(zf + zn) / (zf - zn), -2 * zf * zn / (zf - zn), -1
    hom = proj @ view @ np.vstack([x, np.ones(x.shape[1])])
    sx = (hom[0] / hom[3] * 0.5 + 0.5) * w
    sy = (hom[1] / hom[3] * 0.5 + 0.5) * h
    return sx, sy


def test_frame_3d_matches_orbit_camera():
    from mpm_snow_sim2d_amd.solver import GPUSolver
    P, ps = make_case("ref3d_cubic_coulomb")
    s = GPUSolver(P, ps)
    w, h = 640, 360
    cam = {"angle_deg": 30.0, "elevation_deg": 30.0, "distance": 120.0}
    img = s.render_frame(w, h, camera=cam)
    x = s.get_state()["x"].astype(np.float64)
    sx, sy = _perspective_pixels(x, P, cam, w, h)
    got = _white(img)
    X, Y = np.floor(sx).astype(int), np.floor(sy).astype(int)
    ok = (X >= 0) & (X < w) & (Y >= 0) & (Y < h)
    assert ok.sum() > 0.9 * len(X)
    hit = got[h - 1 - Y[ok], X[ok]]
    # FP32 on the device against FP64 here: a point may land one pixel over
    near = np.zeros_like(hit)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            yy = np.clip(h - 1 - Y[ok] + dy, 0, h - 1)
            xx = np.clip(X[ok] + dx, 0, w - 1)
            near |= got[yy, xx]
    assert near.mean() > 0.999
    assert (img == 128).all(-1).sum() > 100          # the wireframe box
    s.close()


def test_save_frames_loop(tmp_path):
    from mpm_snow_sim2d_amd.solver import GPUSolver, save_frames
    P = snow_params(2, 64, 1e-4)
    P2, ps = make_case("snow2d_quad_sticky")
    s = GPUSolver(P2, ps)
    paths = save_frames(s, str(tmp_path), 3, 10, width=256, height=256)
    assert len(paths) == 3 and all((tmp_path / f"frame_{k}.png").exists() for k in range(3))
    s.close()
    del P
