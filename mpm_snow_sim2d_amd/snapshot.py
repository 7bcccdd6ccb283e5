"""Snapshot files in NumPy (the format of mpm_save_snapshot / mpm_load_snapshot,
include/mpm/mpm.h): read them without a device, or write them from host arrays
(fixtures, a trajectory dataset, resuming from host state).

Layout, little-endian: b"MPMSNAP1", uint32 version (1), uint32 sizeof(mpm_params),
mpm_params, int64 n, int64 substeps, then FP32 x[d][n], v[d][n], C[d*d][n],
F[d*d][n], Jp[n], mass[n], vol[n] in the caller's particle order.
"""
from __future__ import annotations

import ctypes
import struct

import numpy as np

from .params import Params

MAGIC = b"MPMSNAP1"
VERSION = 1
FIELDS = ("x", "v", "C", "F", "Jp", "mass", "vol")


def _shapes(d: int, n: int):
    return {"x": (d, n), "v": (d, n), "C": (d * d, n), "F": (d * d, n), "Jp": (n,), "mass": (n,), "vol": (n,)}


def write_snapshot(path: str, params: Params, state: dict, substeps: int = 0) -> None:
    d = params.dim
    n = int(np.asarray(state["x"]).shape[-1])
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<II", VERSION, ctypes.sizeof(Params)))
        f.write(bytes(params))
        f.write(struct.pack("<qq", n, int(substeps)))
        for k, shp in _shapes(d, n).items():
            a = np.ascontiguousarray(state[k], np.float32)
            if a.size != int(np.prod(shp)):
                raise ValueError(f"{k}: expected {shp}, got {a.shape}")
            f.write(a.astype("<f4").tobytes())


def read_snapshot(path: str) -> tuple[Params, dict, int]:
    with open(path, "rb") as f:
        if f.read(8) != MAGIC:
            raise ValueError("not a snapshot file")
        ver, psz = struct.unpack("<II", f.read(8))
        if ver != VERSION or psz != ctypes.sizeof(Params):
            raise ValueError("unsupported snapshot version or parameter block")
        params = Params.from_buffer_copy(f.read(psz))
        n, substeps = struct.unpack("<qq", f.read(16))
        state = {}
        for k, shp in _shapes(params.dim, n).items():
            cnt = int(np.prod(shp))
            a = np.frombuffer(f.read(4 * cnt), "<f4", count=cnt)
            state[k] = a.astype(np.float32).reshape(shp)
    return params, state, substeps
