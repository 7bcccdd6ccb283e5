"""The seeded scene samplers (mpm_snow_sim2d_amd/scenes.py), CPU.

poisson_disk restates Bridson's sampler as the reference's PoissonDiskCircle /
PoissonDiskSphere use it (src/shared/Shapes.h:17-118, 144-235: 30 attempts per
active point in the annulus [d, 2d), background cells of d / sqrt(DIM), a +-2-cell
neighbourhood test), with a seeded generator instead of std::random_device.  Its
defining properties, checked on disks and balls:

* determinism: the same seed gives the same set, another seed another set;
* separation: every pair of samples is at least min_dist apart;
* containment: every sample lies inside the shape;
* maximality: no point of the shape is farther than 2 min_dist from a sample (an
  active point retires only after 30 candidates in its annulus all failed, so a
  gap that wide would have admitted a candidate; checked on a dense probe set);
* count: the samples cover the shape at the density the spacing implies.

reference_snowballs restates GPUSolver::InitBuffers (GPUSolver.cpp:100-159):
minDist = r / sqrt(NUM_PARTICLES) in 2D, r / cbrt(NUM_PARTICLES) in 3D
(:112,142), per-particle volume 2 pi r^2 / n in 2D (:147, sic: twice the disk's
area) and 4/3 pi r^3 / n in 3D (:117-118), mass = volume * RHO / 100 (:119,148),
all in FP32, the shapes of Parameters.h:43-53.
"""
import math

import numpy as np
import pytest
from scipy.spatial import cKDTree

from mpm_snow_sim2d_amd.scenes import PI_F32, poisson_disk, reference_snowballs


@pytest.mark.parametrize("dim,radius,min_dist", [(2, 15.0, 15.0 / math.sqrt(1500)), (2, 0.08, 0.004),
                                                  (3, 4.0, 4.0 / 800 ** (1 / 3))])
def test_poisson_disk_properties(dim, radius, min_dist):
    c = np.array([50.0, 45.0, 50.0][:dim])
    a = poisson_disk(c, radius, min_dist, np.random.default_rng(7))
    b = poisson_disk(c, radius, min_dist, np.random.default_rng(7))
    other = poisson_disk(c, radius, min_dist, np.random.default_rng(8))
    assert np.array_equal(a, b)
    assert a.shape != other.shape or not np.array_equal(a, other)
    n = a.shape[0]
    assert a.shape == (n, dim) and n > 100
    # containment
    assert (np.linalg.norm(a - c, axis=1) <= radius * (1 + 1e-12)).all()
    # separation (all pairs)
    tree = cKDTree(a)
    d, _ = tree.query(a, k=2)
    assert d[:, 1].min() >= min_dist * (1 - 1e-12)
    # maximality on a dense probe set of the shape
    rng = np.random.default_rng(1)
    probe = rng.uniform(-radius, radius, size=(200_000, dim))
    probe = c + probe[np.linalg.norm(probe, axis=1) <= radius]
    gap, _ = tree.query(probe)
    assert gap.max() <= 2 * min_dist
    # density: a maximal set with separation d covers the shape with one sample per
    # at most the volume of a d/2 ball and at least that of a 2d ball
    vol = math.pi * radius ** 2 if dim == 2 else 4 / 3 * math.pi * radius ** 3
    ball = (lambda r: math.pi * r * r) if dim == 2 else (lambda r: 4 / 3 * math.pi * r ** 3)
    assert vol / ball(2 * min_dist) < n < vol / ball(min_dist / 2)


@pytest.mark.parametrize("dim", [2, 3])
def test_reference_snowballs_volume_and_mass(dim):
    npart = 1000 if dim == 2 else 200
    ps = reference_snowballs(dim, seed=3, num_particles=npart)
    r, rho = np.float32(15.0), np.float32(400.0)
    centers = [np.array([50.0, 45.0, 50.0][:dim]), np.array([150.0, 55.0, 50.0][:dim])]
    x = ps.x.T.astype(np.float64)
    first = np.linalg.norm(x - centers[0], axis=1) <= 15.0 + 1e-4
    second = np.linalg.norm(x - centers[1], axis=1) <= 15.0 + 1e-4
    assert (first ^ second).all()
    for sel, vx in ((first, 40.0), (second, -40.0)):
        n = int(sel.sum())
        if dim == 2:
            vol = np.float32(2.0) * PI_F32 * r * r / np.float32(n)
            md = 15.0 / math.sqrt(npart)
        else:
            vol = (np.float32(4.0) / np.float32(3.0)) * PI_F32 * r * r * r / np.float32(n)
            md = 15.0 / npart ** (1 / 3)
        assert np.all(ps.vol[sel] == vol)
        assert np.all(ps.mass[sel] == (vol * rho) / np.float32(100.0))
        assert np.all(ps.v[0, sel] == vx) and np.all(ps.v[1:, sel] == 0.0)
        d, _ = cKDTree(x[sel]).query(x[sel], k=2)
        assert d[:, 1].min() >= md * (1 - 1e-6)
    # 2D: the reference's volume is twice the disk's area per particle (sic)
    if dim == 2:
        n0 = int(first.sum())
        assert abs(float(ps.vol[first][0]) * n0 - 2 * math.pi * 15.0 ** 2) < 1e-3 * 2 * math.pi * 225
    # seeded: the same seed, the same scene
    again = reference_snowballs(dim, seed=3, num_particles=npart)
    assert np.array_equal(again.x, ps.x)
