"""ctypes front end of the CPU oracle (liboracle: mpm_oracle.c).

TEST INFRASTRUCTURE ONLY — imported by tests/, the smoke check and the CPU-baseline
leg of bench.py, never by the product package.  The oracle restates the reference's
GLSL substep (src/gpu/shaders/{2d,3d}/*.glsl) in FP32 and in an FP64 twin.

State held here follows the reference: APIC matrix B (not C = D^-1 B) and the full
plastic gradient F_P (not its determinant); ``to_solver_state`` converts.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpm_oracle.so")

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def _params_type():
    import sys
    sys.path.insert(0, os.path.dirname(HERE))
    from mpm_snow_sim2d_amd.params import Params
    return Params


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < max(
            os.path.getmtime(os.path.join(HERE, f)) for f in ("mpm_oracle.c", "mpm_oracle_impl.inc", "mpm_oracle.h")):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    Params = _params_type()
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    lib.orc_num_nodes.restype = ctypes.c_long
    lib.orc_num_nodes.argtypes = [P(Params)]
    for s, T in (("f32", ctypes.c_float), ("f64", ctypes.c_double)):
        getattr(lib, f"orc_weight_{s}").restype = T
        getattr(lib, f"orc_weight_{s}").argtypes = [T, ctypes.c_int]
        getattr(lib, f"orc_dweight_{s}").restype = T
        getattr(lib, f"orc_dweight_{s}").argtypes = [T, ctypes.c_int]
        for d in (2, 3):
            f = getattr(lib, f"orc_svd{d}_{s}")
            f.restype = None
            f.argtypes = [vp, vp, vp, vp]
        f = getattr(lib, f"orc_stress_{s}")
        f.restype = None
        f.argtypes = [P(Params), vp, T, vp]
        f = getattr(lib, f"orc_p2g_{s}")
        f.restype = None
        f.argtypes = [P(Params), ctypes.c_long] + [vp] * 10
        f = getattr(lib, f"orc_grid_update_{s}")
        f.restype = None
        f.argtypes = [P(Params)] + [vp] * 5
        f = getattr(lib, f"orc_g2p_{s}")
        f.restype = None
        f.argtypes = [P(Params), ctypes.c_long] + [vp] * 7
        f = getattr(lib, f"orc_run_{s}")
        f.restype = None
        f.argtypes = [P(Params), ctypes.c_long, ctypes.c_int] + [vp] * 11
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _sfx(dtype) -> str:
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


@dataclass
class OracleState:
    x: np.ndarray
    v: np.ndarray
    B: np.ndarray
    Fe: np.ndarray
    Fp: np.ndarray
    mass: np.ndarray
    vol: np.ndarray

    @property
    def dim(self):
        return self.x.shape[0]

    @property
    def n(self):
        return self.x.shape[1]

    @property
    def dtype(self):
        return self.x.dtype

    def copy(self) -> "OracleState":
        return OracleState(*(np.array(a) for a in (self.x, self.v, self.B, self.Fe, self.Fp, self.mass, self.vol)))

    def astype(self, dtype) -> "OracleState":
        return OracleState(*(np.ascontiguousarray(a, dtype) for a in
                             (self.x, self.v, self.B, self.Fe, self.Fp, self.mass, self.vol)))

    @classmethod
    def from_particles(cls, ps, dtype=np.float32) -> "OracleState":
        d, n = ps.dim, ps.n
        eye = np.zeros((d * d, n), dtype)
        for i in range(d):
            eye[i * d + i] = 1
        # always copies: the oracle steps its arrays in place
        return cls(np.array(ps.x, dtype), np.array(ps.v, dtype), np.zeros((d * d, n), dtype),
                   eye.copy(), eye.copy(), np.array(ps.mass, dtype), np.array(ps.vol, dtype))


def dinv(params, dtype=np.float32):
    t = np.dtype(dtype).type
    inv_dx = t(1) / t(params.dx)
    return (t(3) if params.spline_order == 3 else t(4)) * inv_dx * inv_dx


def det_batch(M: np.ndarray, d: int) -> np.ndarray:
    """Determinants of (d*d, n) row-major matrices, same expansion as the oracle."""
    if d == 2:
        return M[0] * M[3] - M[2] * M[1]
    return (M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6])
            + M[2] * (M[3] * M[7] - M[4] * M[6]))


def to_solver_state(params, st: OracleState) -> dict:
    """The oracle's (B, F_P) as the solver stores them: C = D^-1 B, Jp = det F_P."""
    return {"x": st.x, "v": st.v, "C": dinv(params, st.dtype) * st.B, "F": st.Fe, "Jp": det_batch(st.Fp, st.dim)}


def run(params, st: OracleState, nsteps: int) -> OracleState:
    lib = load()
    G = lib.orc_num_nodes(ctypes.byref(params))
    d = st.dim
    gm = np.zeros(G, st.dtype)
    gv = [np.zeros(d * G, st.dtype) for _ in range(3)]
    getattr(lib, f"orc_run_{_sfx(st.dtype)}")(ctypes.byref(params), st.n, int(nsteps), _p(st.x), _p(st.v),
                                              _p(st.mass), _p(st.vol), _p(st.B), _p(st.Fe), _p(st.Fp),
                                              _p(gm), _p(gv[0]), _p(gv[1]), _p(gv[2]))
    return st


def p2g(params, st: OracleState, with_force: bool = False):
    """Grid mass and momentum (and, for the APIC-force scheme, the grid force)."""
    lib = load()
    G = lib.orc_num_nodes(ctypes.byref(params))
    gm = np.zeros(G, st.dtype)
    gmom = np.zeros((st.dim, G), st.dtype)
    gf = np.zeros((st.dim, G), st.dtype)
    getattr(lib, f"orc_p2g_{_sfx(st.dtype)}")(ctypes.byref(params), st.n, _p(st.x), _p(st.v), _p(st.mass),
                                              _p(st.vol), _p(st.B), _p(st.Fe), _p(st.Fp), _p(gm), _p(gmom), _p(gf))
    return (gm, gmom, gf) if with_force else (gm, gmom)


def grid_update(params, gm: np.ndarray, gmom: np.ndarray, gforce: np.ndarray | None = None):
    lib = load()
    vc = np.zeros_like(gmom)
    vf = np.zeros_like(gmom)
    gf = np.zeros_like(gmom) if gforce is None else np.ascontiguousarray(gforce, gmom.dtype)
    getattr(lib, f"orc_grid_update_{_sfx(gm.dtype)}")(ctypes.byref(params), _p(gm), _p(np.ascontiguousarray(gmom)),
                                                      _p(gf), _p(vc), _p(vf))
    return vc, vf


def g2p(params, st: OracleState, vc: np.ndarray, vf: np.ndarray) -> OracleState:
    lib = load()
    getattr(lib, f"orc_g2p_{_sfx(st.dtype)}")(ctypes.byref(params), st.n, _p(st.x), _p(st.v), _p(st.B), _p(st.Fe),
                                              _p(st.Fp), _p(np.ascontiguousarray(vc)), _p(np.ascontiguousarray(vf)))
    return st


def svd(M: np.ndarray):
    """Reference SVD of one matrix (2x2 or 3x3, any float dtype): U, sigma, V."""
    lib = load()
    M = np.ascontiguousarray(M)
    d = M.shape[0]
    U = np.zeros_like(M)
    V = np.zeros_like(M)
    s = np.zeros(d, M.dtype)
    getattr(lib, f"orc_svd{d}_{_sfx(M.dtype)}")(_p(M), _p(U), _p(s), _p(V))
    return U, s, V


def stress(params, F: np.ndarray, Jp: float) -> np.ndarray:
    lib = load()
    F = np.ascontiguousarray(F)
    tau = np.zeros_like(F)
    T = ctypes.c_float if F.dtype == np.float32 else ctypes.c_double
    getattr(lib, f"orc_stress_{_sfx(F.dtype)}")(ctypes.byref(params), _p(F), T(Jp), _p(tau))
    return tau


def weight(x: float, order: int, dtype=np.float64) -> float:
    lib = load()
    return getattr(lib, f"orc_weight_{_sfx(dtype)}")(x, order)


def dweight(x: float, order: int, dtype=np.float64) -> float:
    lib = load()
    return getattr(lib, f"orc_dweight_{_sfx(dtype)}")(x, order)


OMP_LIB_PATH = os.path.join(HERE, "libmpm_oracle_omp.so")
_omp = None


def load_omp():
    """The FP32 substep on every host core (OpenMP build of the same body): the
    all-cores CPU baseline.  Not the checker — its grid sums add in thread order."""
    global _omp
    if _omp is not None:
        return _omp
    if not os.path.exists(OMP_LIB_PATH) or os.path.getmtime(OMP_LIB_PATH) < max(
            os.path.getmtime(os.path.join(HERE, f)) for f in ("mpm_oracle_omp.c", "mpm_oracle_impl.inc")):
        build()
    lib = ctypes.CDLL(OMP_LIB_PATH)
    Params = _params_type()
    vp = ctypes.c_void_p
    lib.orc_num_nodes.restype = ctypes.c_long
    lib.orc_num_nodes.argtypes = [ctypes.POINTER(Params)]
    lib.orc_omp_threads.restype = ctypes.c_int
    lib.orc_omp_threads.argtypes = []
    lib.orc_run_omp_f32.restype = None
    lib.orc_run_omp_f32.argtypes = [ctypes.POINTER(Params), ctypes.c_long, ctypes.c_int] + [vp] * 11
    _omp = lib
    return lib


def omp_threads() -> int:
    return int(load_omp().orc_omp_threads())


def run_omp(params, st: OracleState, nsteps: int) -> OracleState:
    """nsteps FP32 substeps on every host core: the MLS scheme, or the CPU path's
    APIC scheme with the explicit grid force (params.scheme)."""
    if st.dtype != np.float32:
        raise ValueError("the multi-threaded oracle runs in FP32")
    lib = load_omp()
    G = lib.orc_num_nodes(ctypes.byref(params))
    d = st.dim
    gm = np.zeros(G, np.float32)
    gv = [np.zeros(d * G, np.float32) for _ in range(3)]
    lib.orc_run_omp_f32(ctypes.byref(params), st.n, int(nsteps), _p(st.x), _p(st.v), _p(st.mass), _p(st.vol),
                        _p(st.B), _p(st.Fe), _p(st.Fp), _p(gm), _p(gv[0]), _p(gv[1]), _p(gv[2]))
    return st
