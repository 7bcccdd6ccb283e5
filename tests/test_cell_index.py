"""Cell indices of positions (cell_floor / cell_trunc in csrc/mpm_device.hpp, the
base of every stencil, the binning keys and the frame splat).

Within +-2^30 they must be floor / int() truncation exactly (the reference's
floor(x - 0.5) and int(x) stencil bases, shaders/2d/common.glsl); beyond, and for
infinities, they must stay at +-2^30 so that a base plus stencil and tile offsets
never wraps around the int range (a saturating conversion gives INT_MAX, and
INT_MAX + 1 is a negative index); NaN gives 0.  The functions are compiled for
the host from the same header the kernels use (hipcc, no device needed)."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <cstdio>
#include <cmath>
#include <limits>
#include "mpm_device.hpp"

int main()
{
    const float inf = std::numeric_limits<float>::infinity();
    const float vals[] = {0.0f, 0.49f, 0.5f, -0.51f, -1.0f, 255.7f, -3.25f, 1073741760.0f, -1073741760.0f,
                          1073741824.0f, 1.5e9f, 2147483648.0f, -2147483648.0f, 4.0e9f, -4.0e9f, 1e30f,
                          -1e30f, inf, -inf, std::nanf("")};
    for (float v : vals)
        std::printf("%.9g %d %d %d %d\n", v, mpm::cell_floor(v), mpm::cell_trunc(v),
                    mpm::Spline<2>::base(v), mpm::Spline<3>::base(v));
    return 0;
}
"""

LIM = 2 ** 30


def _expect_floor(v):
    if math.isnan(v):
        return 0
    return int(max(min(math.floor(v), LIM), -LIM)) if math.isfinite(v) else (LIM if v > 0 else -LIM)


def _expect_trunc(v):
    if math.isnan(v):
        return 0
    return int(max(min(math.trunc(v), LIM), -LIM)) if math.isfinite(v) else (LIM if v > 0 else -LIM)


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="hipcc not available")
def test_cell_index_exact_in_range_and_bounded_beyond(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    src = tmp_path / "cells.hip"
    src.write_text(PROGRAM)
    exe = tmp_path / "cells"
    cmd = [hipcc, "-std=c++17", "-O2", "--offload-arch=gfx950", "-I",
           os.path.join(ROOT, "mpm_snow_sim2d_amd", "csrc"), str(src), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    rows = [ln.split() for ln in out.stdout.strip().splitlines()]
    assert len(rows) == 20
    for v_s, fl, tr, b2, b3 in rows:
        v = float(np.float32(float(v_s)))
        fl, tr, b2, b3 = int(fl), int(tr), int(b2), int(b3)
        assert fl == _expect_floor(v), (v, fl)
        assert tr == _expect_trunc(v), (v, tr)
        # the stencil bases: floor(x - 0.5) (quadratic) and int(x) - 1 (cubic) in FP32
        assert b2 == _expect_floor(float(np.float32(v) - np.float32(0.5))), (v, b2)
        assert b3 == _expect_trunc(v) - 1, (v, b3)
        # base + offsets of a stencil and of a tile never leave the int range
        for b in (fl, tr, b2, b3):
            assert -2 ** 31 < b - 64 and b + 64 < 2 ** 31
