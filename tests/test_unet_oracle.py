"""Pins the FP64 U-Net oracle (oracle/unet_ref.py) that the device network is
checked against: the convolution against explicit loops, the adjoints of pooling,
upsampling and convolution by inner-product identities, GELU against its
definition, the projection's invariant, BF16 rounding on known bit patterns, and
the whole backward pass against central finite differences.  CPU only."""
import math

import numpy as np
import pytest

from oracle import unet_ref as U


def test_conv3x3_matches_loops():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 5, 6, 3))
    W = rng.normal(size=(9, 4, 3))
    b = rng.normal(size=4)
    y = U.conv3x3(x, W, b)
    ref = np.zeros_like(y)
    for n in range(2):
        for i in range(5):
            for j in range(6):
                for co in range(4):
                    s = b[co]
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            ii, jj = i + dy, j + dx
                            if 0 <= ii < 5 and 0 <= jj < 6:
                                s += W[3 * (dy + 1) + (dx + 1), co] @ x[n, ii, jj]
                    ref[n, i, j, co] = s
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_adjoints():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 8, 6, 5))
    yp = rng.normal(size=(2, 4, 3, 5))
    assert np.isclose((U.pool(x) * yp).sum(), (x * U.pool_bwd(yp)).sum())
    yu = rng.normal(size=(2, 16, 12, 5))
    assert np.isclose((U.up(x) * yu).sum(), (x * U.up_bwd(yu)).sum())
    W = rng.normal(size=(9, 7, 5))
    dz = rng.normal(size=(2, 8, 6, 7))
    assert np.isclose((U.conv3x3(x, W, None) * dz).sum(), (x * U.conv3x3_dgrad(dz, W)).sum())
    assert np.isclose((U.conv3x3(x, W, None) * dz).sum(), (W * U.conv3x3_wgrad(x, dz)).sum())


def test_gelu():
    for z in (-3.0, -1.0, -0.25, 0.0, 0.5, 2.0):
        assert math.isclose(U.gelu(np.float64(z)), 0.5 * z * (1 + math.erf(z / math.sqrt(2))), rel_tol=1e-14,
                            abs_tol=1e-15)
        h = 1e-6
        fd = (U.gelu(np.float64(z + h)) - U.gelu(np.float64(z - h))) / (2 * h)
        assert math.isclose(U.gelu_grad(np.float64(z)), fd, rel_tol=1e-8, abs_tol=1e-9)


def test_bf16_rounding():
    bits = U.to_bf16_bits(np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 0.0], np.float32))
    # 1 + 2^-8 is a tie between 0x3F80 and 0x3F81: even wins; 1 + 3*2^-9 rounds up
    assert list(bits) == [0x3F80, 0x3F80, 0x3F81, 0xC020, 0x0000]
    assert np.array_equal(U.from_bf16_bits(bits)[[0, 3]], np.array([1.0, -2.5], np.float32))


def _random_problem(rng, nb=2, h=8, w=8):
    lay, n = U.layout()
    flat = np.zeros(n)
    for k, (o, s) in lay.items():
        size = int(np.prod(s))
        if k.endswith(".w") and not k.startswith("head"):
            flat[o:o + size] = rng.normal(scale=math.sqrt(2.0 / (9 * s[2])), size=size)
        else:
            flat[o:o + size] = rng.normal(scale=0.1, size=size)
    x = rng.normal(size=(nb, h, w, 32))
    m = rng.uniform(0.0, 1.0, size=(nb, h * w)) * (rng.uniform(size=(nb, h * w)) > 0.3)
    wm = m / m.sum(1, keepdims=True)
    target = rng.normal(scale=0.5, size=(nb, h * w, 2))
    return flat, x, wm, target


def test_projection_removes_weighted_mean():
    rng = np.random.default_rng(2)
    flat, x, wm, _ = _random_problem(rng)
    out = U.forward(U.unpack(flat), x, wm)
    assert np.abs(np.einsum("bi,bio->bo", wm, out)).max() < 1e-12


@pytest.mark.parametrize("block", U.NAMES + ["head"])
def test_backward_matches_finite_differences(block):
    rng = np.random.default_rng(3)
    flat, x, wm, target = _random_problem(rng)
    loss, g = U.loss_and_grad(U.unpack(flat), x, wm, target)
    gflat = U.pack(g)
    lay, _ = U.layout()
    for kind in (".w", ".b"):
        o, s = lay[block + kind]
        size = int(np.prod(s))
        for idx in rng.choice(size, size=min(3, size), replace=False):
            i = o + int(idx)
            h = 1e-5
            fp, fm = flat.copy(), flat.copy()
            fp[i] += h
            fm[i] -= h
            lp, _ = U.loss_and_grad(U.unpack(fp), x, wm, target)
            lm, _ = U.loss_and_grad(U.unpack(fm), x, wm, target)
            fd = (lp - lm) / (2 * h)
            assert math.isclose(gflat[i], fd, rel_tol=1e-5, abs_tol=1e-9 * max(1.0, abs(loss))), (block, kind, idx)
