"""FP64 restatement of the learned surrogate's grid U-Net (forward and backward).

TEST INFRASTRUCTURE ONLY: the checker the device network (libmpm_hip.so,
mpm_surrogate_*) is compared with.  Only tests/, the smoke check and bench.py's
CPU-baseline leg may import it; the product never does.

The network has no reference counterpart (the reference has no learned part); it
follows the parameter layout and architecture the C-ABI documents
(include/mpm/mpm.h, mpm_surrogate_*):

  e0: x(32) -> a0(32)   e1: pool(a0) -> a1(64)   e2: pool(a1) -> a2(128)
  e3: pool(a2) -> a3(128)
  d2: [up(a3) | a2] -> b2(128)   d1: [up(b2) | a1] -> b1(64)   d0: [up(b1) | a0] -> b0(32)
  head: out = b0 @ wh^T + bh, then out' = out - sum_i wm_i out_i per image
  loss = (1/nb) sum_images sum_i wm_i |out'_i - target_i|^2

3x3 convolutions are "same"-padded cross-correlations, tap = 3*(dy+1) + (dx+1),
W[tap][cout][cin]; GELU(z) = z/2 (1 + erf(z/sqrt 2)); pooling is the 2x2 mean,
upsampling the nearest 2x copy.  Tensors are NHWC.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

LAYERS = [  # (c1, c2, cout)
    (32, 0, 32), (32, 0, 64), (64, 0, 128), (128, 0, 128),
    (128, 128, 128), (128, 64, 64), (64, 32, 32),
]
NAMES = ["e0", "e1", "e2", "e3", "d2", "d1", "d0"]


def layout():
    """Offsets of every parameter block in the flat FP32 vector."""
    off = 0
    out = {}
    for name, (c1, c2, co) in zip(NAMES, LAYERS):
        n = 9 * co * (c1 + c2)
        out[name + ".w"] = (off, (9, co, c1 + c2))
        off += n
        out[name + ".b"] = (off, (co,))
        off += co
    out["head.w"] = (off, (2, 32))
    off += 64
    out["head.b"] = (off, (2,))
    off += 2
    return out, off


def unpack(flat: np.ndarray) -> dict:
    lay, n = layout()
    assert flat.size == n, (flat.size, n)
    return {k: flat[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in lay.items()}


def pack(p: dict) -> np.ndarray:
    lay, n = layout()
    flat = np.zeros(n, np.float64)
    for k, (o, s) in lay.items():
        flat[o:o + int(np.prod(s))] = np.asarray(p[k]).reshape(-1)
    return flat


# ----------------------------------------------------------------------------- bf16
def to_bf16_bits(x: np.ndarray) -> np.ndarray:
    """Round FP32 to BF16 (nearest, ties to even); returns the 16-bit patterns."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    return ((u + r) >> 16).astype(np.uint16)


def from_bf16_bits(b: np.ndarray) -> np.ndarray:
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


def bf16_round(x: np.ndarray) -> np.ndarray:
    return from_bf16_bits(to_bf16_bits(x)).astype(np.float64)


# ----------------------------------------------------------------------------- ops
def _shift(x: np.ndarray, dy: int, dx: int) -> np.ndarray:
    """s[:, y, x] = x[:, y+dy, x+dx] with zero outside."""
    nb, h, w, c = x.shape
    s = np.zeros_like(x)
    ys, yd = (slice(dy, h), slice(0, h - dy)) if dy >= 0 else (slice(0, h + dy), slice(-dy, h))
    xs, xd = (slice(dx, w), slice(0, w - dx)) if dx >= 0 else (slice(0, w + dx), slice(-dx, w))
    s[:, yd, xd] = x[:, ys, xs]
    return s


def conv3x3(x: np.ndarray, W: np.ndarray, b: np.ndarray | None) -> np.ndarray:
    nb, h, w, _ = x.shape
    y = np.zeros((nb, h, w, W.shape[1]))
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        y += _shift(x, dy, dx) @ W[tap].T
    if b is not None:
        y += b
    return y


def conv3x3_wgrad(x: np.ndarray, dz: np.ndarray) -> np.ndarray:
    cin, cout = x.shape[-1], dz.shape[-1]
    dW = np.zeros((9, cout, cin))
    dzf = dz.reshape(-1, cout)
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        dW[tap] = dzf.T @ _shift(x, dy, dx).reshape(-1, cin)
    return dW


def conv3x3_dgrad(dz: np.ndarray, W: np.ndarray) -> np.ndarray:
    nb, h, w, _ = dz.shape
    dx_ = np.zeros((nb, h, w, W.shape[2]))
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        dx_ += _shift(dz @ W[tap], -dy, -dx)
    return dx_


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def pool(x):
    nb, h, w, c = x.shape
    return x.reshape(nb, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def pool_bwd(dy):
    return np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2) / 4.0


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_bwd(dy):
    nb, h, w, c = dy.shape
    return dy.reshape(nb, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))


# ----------------------------------------------------------------------------- network
def forward(p: dict, x: np.ndarray, wm: np.ndarray, cache: dict | None = None) -> np.ndarray:
    """x: [nb, h, w, 32]; wm: [nb, h*w]; returns the projected output [nb, h*w, 2]."""
    c = {} if cache is None else cache

    def layer(name, inp):
        z = conv3x3(inp, p[name + ".w"], p[name + ".b"])
        c[name + ".in"], c[name + ".z"] = inp, z
        return gelu(z)

    a0 = layer("e0", x)
    a1 = layer("e1", pool(a0))
    a2 = layer("e2", pool(a1))
    a3 = layer("e3", pool(a2))
    b2 = layer("d2", np.concatenate([up(a3), a2], -1))
    b1 = layer("d1", np.concatenate([up(b2), a1], -1))
    b0 = layer("d0", np.concatenate([up(b1), a0], -1))
    nb = x.shape[0]
    feat = b0.reshape(nb, -1, 32)
    out = feat @ p["head.w"].T + p["head.b"]
    c["b0"] = feat
    s = np.einsum("bi,bio->bo", wm, out)
    return out - s[:, None, :]


def loss_and_grad(p: dict, x: np.ndarray, wm: np.ndarray, target: np.ndarray):
    c: dict = {}
    out = forward(p, x, wm, c)
    nb = x.shape[0]
    d = out - target
    loss = float(np.einsum("bi,bio->", wm, d * d) / nb)
    g = {}
    dproj = 2.0 * wm[..., None] * d / nb
    dout = dproj - wm[..., None] * dproj.sum(1, keepdims=True)
    g["head.w"] = np.einsum("bio,bic->oc", dout, c["b0"])
    g["head.b"] = dout.sum((0, 1))
    h, w = x.shape[1], x.shape[2]
    db0 = (dout @ p["head.w"]).reshape(nb, h, w, 32)

    def back(name, da):
        dz = da * gelu_grad(c[name + ".z"])
        g[name + ".w"] = conv3x3_wgrad(c[name + ".in"], dz)
        g[name + ".b"] = dz.sum((0, 1, 2))
        return conv3x3_dgrad(dz, p[name + ".w"])

    dcat = back("d0", db0)
    db1, da0 = up_bwd(dcat[..., :64]), dcat[..., 64:]
    dcat = back("d1", db1)
    db2, da1 = up_bwd(dcat[..., :128]), dcat[..., 128:]
    dcat = back("d2", db2)
    da3, da2 = up_bwd(dcat[..., :128]), dcat[..., 128:]
    da2 = da2 + pool_bwd(back("e3", da3))
    da1 = da1 + pool_bwd(back("e2", da2))
    da0 = da0 + pool_bwd(back("e1", da1))
    back("e0", da0)
    return loss, g
