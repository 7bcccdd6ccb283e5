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
