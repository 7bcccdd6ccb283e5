"""The C-ABI library loads and exports every entry point include/mpm/*.h declares;
host-only entry points behave (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from mpm_snow_sim2d_amd import _lib
from mpm_snow_sim2d_amd.params import COptions, Params, reference_params, snow_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for hdr in ("mpm.h", "params.h"):
        text = open(os.path.join(ROOT, "include", "mpm", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(mpm_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    decl = _declared()
    assert decl, "no declarations found"
    missing = [n for n in sorted(decl) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.EXPORTS) == decl


def test_params_reference_matches_parameters_h():
    lib = _lib.load()
    for dim in (2, 3):
        p = Params()
        assert lib.mpm_params_reference(dim, ctypes.byref(p)) == 0
        want = reference_params(dim)
        for name, _ in Params._fields_:
            assert getattr(p, name) == pytest.approx(getattr(want, name), rel=0, abs=0), name
    # MU and LAMBDA in FP32 exactly as Parameters.h:82-84
    assert np.float32(p.MU) == np.float32(np.float32(1.4e5) / (np.float32(2) * (np.float32(1) + np.float32(0.2))))
    assert abs(p.MU - 58333.33) < 0.01 and abs(p.LAMBDA - 38888.89) < 0.01


def test_invalid_arguments_fail_cleanly():
    lib = _lib.load()
    p = reference_params(2)
    p.dim = 4
    h = ctypes.c_void_p()
    code = lib.mpm_create(ctypes.byref(p), None, None, ctypes.byref(h))
    assert code == _lib.MPM_E_INVALID
    assert b"dim" in lib.mpm_last_error()
    assert lib.mpm_step(None, 1) == _lib.MPM_E_INVALID
    assert lib.mpm_sort_pairs_host(None, 5, 40, None, None) == _lib.MPM_E_INVALID
    o = COptions()
    assert lib.mpm_options_default(ctypes.byref(o)) == 0 and o.fused == 1


def test_num_nodes():
    lib = _lib.load()
    n = ctypes.c_int64()
    p = snow_params(3, 256, 5e-5)
    assert lib.mpm_num_nodes(ctypes.byref(p), ctypes.byref(n)) == 0
    assert n.value == 257 ** 3 == p.num_nodes


def _cparts(n, dim):
    x = np.full((dim, max(n, 1)), 0.5, np.float32)
    m = np.ones(max(n, 1), np.float32)
    keep = [x, m]
    cp = _lib.CParticles(n, x.ctypes.data, None, m.ctypes.data, m.ctypes.data, None, None, None)
    return cp, keep


def test_batch_create_validates_scenes():
    """Scene sets a batch cannot run are refused before any device work."""
    lib = _lib.load()
    h = ctypes.c_void_p()
    a, b = snow_params(2, 32, 1e-4), snow_params(3, 16, 1e-4)
    ca, ka = _cparts(10, 2)
    cb, kb = _cparts(10, 3)
    P = (Params * 2)(a, b)
    C = (_lib.CParticles * 2)(ca, cb)
    assert lib.mpm_batch_create(2, P, C, None, ctypes.byref(h)) == _lib.MPM_E_INVALID
    assert "share" in lib.mpm_last_error().decode()
    assert lib.mpm_batch_create(0, P, C, None, ctypes.byref(h)) == _lib.MPM_E_INVALID
    c = snow_params(2, 32, 1e-4)
    c.spline_order = 3
    P2 = (Params * 2)(a, c)
    C2 = (_lib.CParticles * 2)(ca, ca)
    assert lib.mpm_batch_create(2, P2, C2, None, ctypes.byref(h)) == _lib.MPM_E_INVALID
    d = snow_params(2, 32, 1e-4)
    d.scheme = 1
    P3 = (Params * 2)(d, d)
    assert lib.mpm_batch_create(2, P3, C2, None, ctypes.byref(h)) == _lib.MPM_E_INVALID
    assert "MLS" in lib.mpm_last_error().decode()
    assert not h.value
    best = COptions()
    assert lib.mpm_autotune(2, P, C, 0, 1, ctypes.byref(best), None) == _lib.MPM_E_INVALID


def test_batch_scene_generator():
    from mpm_snow_sim2d_amd.scenes import batch_scenes
    a = batch_scenes(16, seed=3)
    b = batch_scenes(16, seed=3)
    assert len(a) == 16
    for (Pa, pa), (Pb, pb) in zip(a, b):
        assert np.array_equal(pa.x, pb.x) and Pa.WIDTH == Pb.WIDTH
        n = pa.n
        assert 0.8 * 5_000 <= n <= 1.2 * 80_000
        assert Pa.WIDTH == (128 if pa.meta["n_target"] < 20_000 else 256)
        assert 1 <= pa.meta["shapes"] <= 3
        assert pa.x.min() > 0.05 and pa.x.max() < 0.95
