"""Seeded initial particle sets.

The reference samples its snowballs with Bridson's Poisson-disk algorithm from an
unseeded ``std::random_device`` (src/shared/Shapes.h:17-118, 144-235), so no run is
reproducible.  Here every sampler takes a seed:

* ``poisson_disk``   Bridson (2007) in 2D/3D with the reference's constants:
                      background cell minDist/sqrt(DIM), 30 attempts per active point,
                      candidates in the annulus [minDist, 2 minDist), neighbour check
                      over +-2 cells (Shapes.h:144-235 and :17-118);
* ``jitter_lattice`` a lattice of spacing h, each point jittered by U(-h/2, h/2) per
                      axis, kept inside the shape — what the benchmark scenes use.

Scenes return a ``ParticleSet`` (FP32, component-major ``x[d, p]``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

PI_F32 = np.float32(3.1415926)   # Shapes.h:8


@dataclass
class ParticleSet:
    x: np.ndarray                     # (dim, n) float32
    v: np.ndarray                     # (dim, n) float32
    mass: np.ndarray                  # (n,) float32
    vol: np.ndarray                   # (n,) float32
    meta: dict = field(default_factory=dict)

    @property
    def dim(self) -> int:
        return self.x.shape[0]

    @property
    def n(self) -> int:
        return self.x.shape[1]

    def concat(self, other: "ParticleSet") -> "ParticleSet":
        return ParticleSet(np.concatenate([self.x, other.x], 1), np.concatenate([self.v, other.v], 1),
                           np.concatenate([self.mass, other.mass]), np.concatenate([self.vol, other.vol]),
                           dict(self.meta))

    @staticmethod
    def empty(dim: int) -> "ParticleSet":
        z = np.zeros((dim, 0), np.float32)
        return ParticleSet(z, z.copy(), np.zeros(0, np.float32), np.zeros(0, np.float32))


# ----------------------------------------------------------------------------- samplers
def poisson_disk(center, radius: float, min_dist: float, rng: np.random.Generator,
                 max_attempts: int = 30) -> np.ndarray:
    """Bridson Poisson-disk samples inside a disk (2D) or ball (3D); returns (n, dim)."""
    c = np.asarray(center, np.float64)
    dim = c.shape[0]
    cell = min_dist / math.sqrt(dim)
    gw = int(math.ceil(2.0 * radius / cell)) + 2
    grid = -np.ones((gw,) * dim, np.int64)
    origin = c - radius
    samples: list[np.ndarray] = []
    active: list[int] = []

    def to_cell(p):
        return tuple(int((p[a] - origin[a]) / cell) for a in range(dim))

    def add(p):
        grid[to_cell(p)] = len(samples)
        samples.append(p)
        active.append(len(samples) - 1)

    # seed: uniform point in the shape (Shapes.h:196-200, 3D :53-63)
    if dim == 2:
        a = rng.uniform(0, 2 * math.pi)
        r = math.sqrt(rng.uniform(0, radius * radius))
        add(c + np.array([r * math.cos(a), r * math.sin(a)]))
    else:
        th = rng.uniform(0, 2 * math.pi)
        ph = math.acos(rng.uniform(-1, 1))
        r = rng.uniform(0, radius ** 3) ** (1.0 / 3.0)
        add(c + r * np.array([math.sin(ph) * math.cos(th), math.sin(ph) * math.sin(th), math.cos(ph)]))

    offs = np.stack(np.meshgrid(*([np.arange(-2, 3)] * dim), indexing="ij"), -1).reshape(-1, dim)
    md2 = min_dist * min_dist
    while active:
        ai = int(rng.integers(0, len(active)))
        point = samples[active[ai]]
        found = False
        for _ in range(max_attempts):
            d = min_dist * (rng.uniform(0, 1) + 1.0)
            if dim == 2:
                a = rng.uniform(0, 2 * math.pi)
                cand = point + d * np.array([math.cos(a), math.sin(a)])
            else:
                th = rng.uniform(0, 2 * math.pi)
                ph = math.acos(rng.uniform(-1, 1))
                cand = point + d * np.array([math.sin(ph) * math.cos(th), math.sin(ph) * math.sin(th),
                                             math.cos(ph)])
            if np.sum((cand - c) ** 2) > radius * radius:
                continue
            cc = np.array(to_cell(cand))
            if (cc < 0).any() or (cc >= gw).any():
                continue
            nb = cc + offs
            nb = nb[((nb >= 0) & (nb < gw)).all(1)]
            ids = grid[tuple(nb.T)]
            ids = ids[ids >= 0]
            if ids.size and (np.sum((np.asarray([samples[i] for i in ids]) - cand) ** 2, 1) < md2).any():
                continue
            add(cand)
            found = True
            break
        if not found:
            active.pop(ai)
    return np.asarray(samples, np.float64)


def jitter_lattice(lo, hi, h: float, rng: np.random.Generator, inside=None) -> np.ndarray:
    """Lattice points of spacing h covering [lo, hi), jittered by U(-h/2, h/2); (n, dim)."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    counts = np.maximum(np.floor((hi - lo) / h + 1e-9).astype(int), 1)
    axes = [lo[a] + (np.arange(counts[a]) + 0.5) * h for a in range(lo.shape[0])]
    pts = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, lo.shape[0])
    pts = pts + rng.uniform(-0.5 * h, 0.5 * h, size=pts.shape)
    if inside is not None:
        pts = pts[inside(pts)]
    return pts


def _set(points: np.ndarray, vel, vol: float, mass: float) -> ParticleSet:
    n = points.shape[0]
    dim = points.shape[1] if n else len(vel)
    x = np.ascontiguousarray(points.T.astype(np.float32)).reshape(dim, n)
    v = np.repeat(np.asarray(vel, np.float32)[:, None], n, 1)
    return ParticleSet(x, np.ascontiguousarray(v), np.full(n, mass, np.float32), np.full(n, vol, np.float32))


# ----------------------------------------------------------------------------- scenes
def reference_snowballs(dim: int = 2, seed: int = 0, num_particles: int = 3000) -> ParticleSet:
    """The reference's default scene (Parameters.h:43-53): snowballs of radius 15 at
    (50,45[,50]) moving +40 and at (150,55[,50]) moving -40, grid units.  minDist =
    r/sqrt(NUM_PARTICLES) in 2D, r/cbrt(NUM_PARTICLES) in 3D; per-particle volume
    2 pi r^2 / n (2D, as GPUSolver.cpp:147 has it) or 4/3 pi r^3 / n (3D), mass =
    vol * RHO / 100 (GPUSolver.cpp:119,148)."""
    rng = np.random.default_rng(seed)
    balls = [((50.0, 45.0, 50.0), (40.0, 0.0, 0.0)), ((150.0, 55.0, 50.0), (-40.0, 0.0, 0.0))]
    out = ParticleSet.empty(dim)
    r = np.float32(15.0)
    rho = np.float32(400.0)
    for c, v in balls:
        if dim == 2:
            md = float(r / np.sqrt(np.float32(num_particles)))
        else:
            md = float(r / np.cbrt(np.float32(num_particles)))
        pts = poisson_disk(c[:dim], float(r), md, rng)
        n = np.float32(pts.shape[0])
        if dim == 2:
            vol = np.float32(2.0) * PI_F32 * r * r / n
        else:
            vol = (np.float32(4.0) / np.float32(3.0)) * PI_F32 * r * r * r / n
        mass = (vol * rho) / np.float32(100.0)
        out = out.concat(_set(pts, v[:dim], float(vol), float(mass)))
    out.meta = {"scene": "reference_snowballs", "seed": seed, "dim": dim}
    return out


def snowball_disks(seed: int = 0, cells: int = 128, per_cell: float = 30.4, rho: float = 400.0) -> ParticleSet:
    """Plain 2D snow: disks r=0.08 at (0.3,0.5) and (0.7,0.55) in the unit square with
    velocities (+3,0) and (-3,0), jitter-sampled at `per_cell` particles per cell
    (30.4 gives ~20k particles on a 128^2 grid)."""
    rng = np.random.default_rng(seed)
    dx = 1.0 / cells
    h = dx / math.sqrt(per_cell)
    out = ParticleSet.empty(2)
    for c, v in (((0.3, 0.5), (3.0, 0.0)), ((0.7, 0.55), (-3.0, 0.0))):
        c = np.asarray(c)
        pts = jitter_lattice(c - 0.08, c + 0.08, h, rng, lambda p, c=c: ((p - c) ** 2).sum(1) <= 0.08 ** 2)
        out = out.concat(_set(pts, v, h * h, rho * h * h))
    out.meta = {"scene": "snowball_disks", "seed": seed, "h": h}
    return out


def falling_cube(seed: int = 0, side: float = 0.3, n_side: int = 126, bottom: float = 0.5,
                 rho: float = 400.0) -> ParticleSet:
    """Plain 3D snow: a cube of side `side` m on an n_side^3 jittered lattice
    (126^3 = 2,000,376 particles), centred at x = z = 0.5 with its lower face at
    y = `bottom`, at rest."""
    rng = np.random.default_rng(seed)
    h = side / n_side
    lo = np.array([0.5 - side / 2, bottom, 0.5 - side / 2])
    axes = [lo[a] + (np.arange(n_side, dtype=np.float64) + 0.5) * h for a in range(3)]
    n = n_side ** 3
    x = np.empty((3, n), np.float32)
    # build component by component to keep peak memory low at 2M points
    grids = np.meshgrid(*axes, indexing="ij")
    for a in range(3):
        x[a] = (grids[a].reshape(-1) + rng.uniform(-0.5 * h, 0.5 * h, n)).astype(np.float32)
    del grids
    v = np.zeros((3, n), np.float32)
    vol = h ** 3
    ps = ParticleSet(x, v, np.full(n, rho * vol, np.float32), np.full(n, vol, np.float32))
    ps.meta = {"scene": "falling_cube", "seed": seed, "h": h, "n_side": n_side}
    return ps


def colliding_spheres(seed: int = 0, n_target: int = 16_000_000, count: int = 4, rho: float = 400.0,
                      per_cell_dx: float = 1.0 / 512) -> ParticleSet:
    """Plain 3D large: `count` spheres of radius U(0.08, 0.15) at seeded
    non-overlapping centres in [0.2, 0.8]^3, each moving towards the spheres'
    common centroid at U(2, 4) m/s plus a U(-0.5, 0.5) m/s perturbation per
    component (components within +-4 m/s), so that they collide; jitter-sampled so
    the total is about n_target particles.  meta holds the centres, radii and
    velocities.  Seed 2 (LARGE_SEED) brings the first two spheres into contact
    after ~157 substeps of 2.5e-5 s and three pairs within 1000, all far from the
    walls."""
    rng = np.random.default_rng(seed)
    spheres = []
    while len(spheres) < count:
        r = rng.uniform(0.08, 0.15)
        c = rng.uniform(0.2, 0.8, 3)
        if all(np.linalg.norm(c - c2) > r + r2 + 0.01 for c2, r2 in spheres):
            spheres.append((c, r))
    mid = np.mean([c for c, _ in spheres], axis=0)
    vels = []
    for c, _ in spheres:
        d = (mid - c) / max(float(np.linalg.norm(mid - c)), 1e-9)
        vels.append(np.clip(rng.uniform(2.0, 4.0) * d + rng.uniform(-0.5, 0.5, 3), -4.0, 4.0))
    total_vol = sum(4.0 / 3.0 * math.pi * r ** 3 for _, r in spheres)
    h = (total_vol / n_target) ** (1.0 / 3.0)
    out = ParticleSet.empty(3)
    for (c, r), v in zip(spheres, vels):
        pts = jitter_lattice(c - r, c + r, h, rng, lambda p, c=c, r=r: ((p - c) ** 2).sum(1) <= r * r)
        out = out.concat(_set(pts, v, h ** 3, rho * h ** 3))
    out.meta = {"scene": "colliding_spheres", "seed": seed, "h": h,
                "centres": [c.tolist() for c, _ in spheres], "radii": [float(r) for _, r in spheres],
                "velocities": [v.tolist() for v in vels]}
    return out


# the large workload's seed: its spheres meet within the configured 1000 substeps
LARGE_SEED = 2


def random_scene(seed: int, n_target: int, rho: float = 400.0, max_shapes: int = 3) -> ParticleSet:
    """Random 2D snow scene in the unit square: 1..max_shapes disks (r ~ U(0.05, 0.12))
    or boxes (half-sides ~ U(0.04, 0.10)) at seeded non-overlapping centres in
    [0.2, 0.8]^2, each with its own velocity U(-3, 3) m/s per axis, jitter-sampled on
    one lattice whose spacing puts about n_target particles in the union."""
    rng = np.random.default_rng(seed)
    k = int(rng.integers(1, max_shapes + 1))
    shapes = []
    attempts = 0
    while len(shapes) < k and attempts < 1000:
        attempts += 1
        box = bool(rng.integers(0, 2))
        c = rng.uniform(0.2, 0.8, 2)
        half = rng.uniform(0.04, 0.10, 2) if box else np.full(2, rng.uniform(0.05, 0.12))
        reach = float(np.hypot(*half)) if box else float(half[0])
        if all(np.linalg.norm(c - c2) > reach + r2 + 0.01 for _, c2, _, r2 in shapes):
            shapes.append((box, c, half, reach))
    area = sum(4.0 * h[0] * h[1] if box else math.pi * h[0] ** 2 for box, _, h, _ in shapes)
    h = math.sqrt(area / n_target)
    out = ParticleSet.empty(2)
    for box, c, half, _ in shapes:
        if box:
            inside = lambda p, c=c, half=half: (np.abs(p - c) <= half).all(1)
        else:
            inside = lambda p, c=c, r=half[0]: ((p - c) ** 2).sum(1) <= r * r
        pts = jitter_lattice(c - half, c + half, h, rng, inside)
        out = out.concat(_set(pts, rng.uniform(-3.0, 3.0, 2), h * h, rho * h * h))
    out.meta = {"scene": "random_scene", "seed": seed, "h": h, "shapes": len(shapes), "n_target": n_target}
    return out


def batch_scenes(nscenes: int = 64, seed: int = 0, n_min: int = 5_000, n_max: int = 80_000,
                 split: int = 20_000):
    """Batched 2D workload: `nscenes` independent random scenes (``random_scene``)
    with particle counts log-uniform in [n_min, n_max]; a scene under `split`
    particles gets a 128^2 grid (dt 1e-4), a larger one 256^2 (dt 5e-5).  Returns a
    list of (Params, ParticleSet)."""
    from .params import snow_params
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(math.log(n_min), math.log(n_max), nscenes)).astype(int)
    out = []
    for i, n in enumerate(counts):
        cells, dt = (128, 1e-4) if n < split else (256, 5e-5)
        out.append((snow_params(2, cells, dt), random_scene(seed * 100_003 + i, int(n))))
    return out


def colliding_scene(seed: int, n_target: int, rho: float = 400.0, max_shapes: int = 3,
                    speed=(1.0, 3.0)) -> ParticleSet:
    """Random 2D scene of interacting bodies: 2..max_shapes disks (r ~ U(0.05, 0.10))
    or boxes (half-sides ~ U(0.04, 0.08)) at seeded non-overlapping centres in
    [0.25, 0.75]^2, each moving towards the middle of the square at U(speed) m/s
    plus a U(-0.5, 0.5) m/s perturbation per axis; one jittered lattice of about
    n_target particles in the union."""
    rng = np.random.default_rng(seed)
    k = int(rng.integers(2, max_shapes + 1))
    shapes = []
    attempts = 0
    while len(shapes) < k and attempts < 1000:
        attempts += 1
        box = bool(rng.integers(0, 2))
        c = rng.uniform(0.25, 0.75, 2)
        half = rng.uniform(0.04, 0.08, 2) if box else np.full(2, rng.uniform(0.05, 0.10))
        reach = float(np.hypot(*half)) if box else float(half[0])
        if all(np.linalg.norm(c - c2) > reach + r2 + 0.005 for _, c2, _, r2 in shapes):
            shapes.append((box, c, half, reach))
    area = sum(4.0 * h[0] * h[1] if box else math.pi * h[0] ** 2 for box, _, h, _ in shapes)
    h = math.sqrt(area / n_target)
    out = ParticleSet.empty(2)
    mid = np.array([0.5, 0.5])
    for box, c, half, _ in shapes:
        if box:
            inside = lambda p, c=c, half=half: (np.abs(p - c) <= half).all(1)
        else:
            inside = lambda p, c=c, r=half[0]: ((p - c) ** 2).sum(1) <= r * r
        pts = jitter_lattice(c - half, c + half, h, rng, inside)
        d = mid - c
        d = d / max(np.linalg.norm(d), 1e-9)
        vel = d * rng.uniform(*speed) + rng.uniform(-0.5, 0.5, 2)
        out = out.concat(_set(pts, vel, h * h, rho * h * h))
    out.meta = {"scene": "colliding_scene", "seed": seed, "h": h, "shapes": len(shapes), "n_target": n_target}
    return out
