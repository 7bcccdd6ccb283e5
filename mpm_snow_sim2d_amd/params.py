"""Run-time parameter set, a ctypes mirror of ``mpm_params`` (include/mpm/params.h).

Field names follow the reference's ``src/shared/Parameters.h`` (WIDTH, HEIGHT, DEPTH,
BORDER_MARGIN, GRAVITY, FRICTION, DELTA_TIME, THETA_C, THETA_S, XI, NU, E, RHO, MU,
LAMBDA); the lower-case fields are what the reference fixes in code (dimension,
spline order, boundary kind, grid spacing, light-node threshold).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

BOUNDARY_COULOMB_MARGIN = 0
BOUNDARY_STICKY = 1
SCHEME_MLS = 0          # the GPU path (GLSL shaders)
SCHEME_APIC_FORCE = 1   # the CPU path (src/cpu): APIC + explicit grid force


class Params(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("WIDTH", ctypes.c_int),
        ("HEIGHT", ctypes.c_int),
        ("DEPTH", ctypes.c_int),
        ("spline_order", ctypes.c_int),
        ("boundary", ctypes.c_int),
        ("BORDER_MARGIN", ctypes.c_float),
        ("dx", ctypes.c_float),
        ("GRAVITY", ctypes.c_float),
        ("FRICTION", ctypes.c_float),
        ("DELTA_TIME", ctypes.c_float),
        ("THETA_C", ctypes.c_float),
        ("THETA_S", ctypes.c_float),
        ("XI", ctypes.c_float),
        ("NU", ctypes.c_float),
        ("E", ctypes.c_float),
        ("RHO", ctypes.c_float),
        ("MU", ctypes.c_float),
        ("LAMBDA", ctypes.c_float),
        ("mass_eps", ctypes.c_float),
        ("scheme", ctypes.c_int),
    ]

    def derive_lame(self) -> "Params":
        """MU, LAMBDA from E, NU in FP32, as Parameters.h:82-84 computes them."""
        E, NU = np.float32(self.E), np.float32(self.NU)
        one, two = np.float32(1.0), np.float32(2.0)
        self.MU = float(E / (two * (one + NU)))
        self.LAMBDA = float((E * NU) / ((one + NU) * (one - (two * NU))))
        return self

    @property
    def num_nodes(self) -> int:
        g = (self.WIDTH + 1) * (self.HEIGHT + 1)
        return g * (self.DEPTH + 1) if self.dim == 3 else g

    def copy(self) -> "Params":
        p = Params()
        ctypes.pointer(p)[0] = self
        return p

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}

    def __repr__(self) -> str:  # pragma: no cover - debugging aid
        return "Params(" + ", ".join(f"{k}={v!r}" for k, v in self.as_dict().items()) + ")"


def reference_params(dim: int = 2) -> Params:
    """The reference's Parameters.h values (grid units, cubic spline, Coulomb walls),
    with the GPU path's MLS scheme; ``scheme=SCHEME_APIC_FORCE`` selects the CPU path's."""
    p = Params(
        dim=dim, WIDTH=200, HEIGHT=100, DEPTH=100,            # Parameters.h:30-36
        spline_order=3, boundary=BOUNDARY_COULOMB_MARGIN,
        BORDER_MARGIN=2.0, dx=1.0,
        GRAVITY=-9.81, FRICTION=0.3,                          # Parameters.h:58-60
        DELTA_TIME=1e-3,                                      # Parameters.h:68
        THETA_C=2.0e-2, THETA_S=6.0e-3, XI=10.0,              # Parameters.h:74-76
        NU=0.2, E=1.4e5, RHO=400.0,                           # Parameters.h:77-79
        mass_eps=1e-6,                                        # shaders/2d/grid_update.glsl:90
    )
    return p.derive_lame()


def snow_params(dim: int, cells: int, dt: float, spline_order: int = 2,
                boundary: int = BOUNDARY_STICKY, margin: float = 3.0) -> Params:
    """North-star configuration: unit domain of `cells` cells per axis in metres,
    quadratic MLS, sticky walls, Stomakhin parameters E=1.4e5, nu=0.2,
    theta_c=2.5e-2, theta_s=7.5e-3, xi=10, rho=400, g=-9.8."""
    p = Params(
        dim=dim, WIDTH=cells, HEIGHT=cells, DEPTH=cells if dim == 3 else 0,
        spline_order=spline_order, boundary=boundary,
        BORDER_MARGIN=margin, dx=1.0 / cells,
        GRAVITY=-9.8, FRICTION=0.3, DELTA_TIME=dt,
        THETA_C=2.5e-2, THETA_S=7.5e-3, XI=10.0,
        NU=0.2, E=1.4e5, RHO=400.0,
        mass_eps=0.0,
    )
    return p.derive_lame()


@dataclass
class Options:
    """Solver options beyond the physics (``mpm_options``): binning period; pipeline
    (0 the four reference stages, 1 fused over one-block work items, 2 / 3 fused
    over 2x2x2-block work items with a drift margin of 0.5 / 1.5 cells, one 3D scene
    only); particles per work item at most (0: the pipeline's maximum)."""
    sort_every: int = 16
    fused: int = 1
    tile_particles: int = 512


class COptions(ctypes.Structure):
    _fields_ = [("sort_every", ctypes.c_int), ("fused", ctypes.c_int), ("tile_particles", ctypes.c_int)]

    @classmethod
    def of(cls, o: Options) -> "COptions":
        return cls(o.sort_every, o.fused, o.tile_particles)
