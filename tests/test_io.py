"""Snapshot format and PNG writer on the host (no device): the NumPy writer and the
library's header reader agree, the NumPy reader round-trips, bad files are refused,
and mpm_write_png produces a PNG that zlib decodes to the same pixels."""
import ctypes
import struct
import zlib

import numpy as np

from mpm_snow_sim2d_amd import _lib
from mpm_snow_sim2d_amd.params import Params, reference_params, snow_params
from mpm_snow_sim2d_amd.snapshot import read_snapshot, write_snapshot


def _state(d, n, rng):
    return {"x": rng.uniform(size=(d, n)), "v": rng.normal(size=(d, n)), "C": rng.normal(size=(d * d, n)),
            "F": rng.normal(size=(d * d, n)), "Jp": rng.uniform(0.9, 1.1, n), "mass": rng.uniform(size=n),
            "vol": rng.uniform(size=n)}


def test_snapshot_numpy_roundtrip_and_library_header(tmp_path):
    rng = np.random.default_rng(0)
    for P in (reference_params(2), snow_params(3, 32, 1e-4)):
        st = _state(P.dim, 17, rng)
        path = str(tmp_path / f"s{P.dim}.snap")
        write_snapshot(path, P, st, substeps=123)
        P2, st2, steps = read_snapshot(path)
        assert steps == 123 and bytes(P2) == bytes(P)
        for k in st:
            assert np.array_equal(st2[k], np.asarray(st[k], np.float32).reshape(st2[k].shape))
        lib = _lib.load()
        P3 = Params()
        n = ctypes.c_int64(0)
        s = ctypes.c_int64(0)
        assert lib.mpm_snapshot_info(path.encode(), ctypes.byref(P3), ctypes.byref(n), ctypes.byref(s)) == 0
        assert bytes(P3) == bytes(P) and n.value == 17 and s.value == 123


def test_snapshot_rejects_bad_files(tmp_path):
    lib = _lib.load()
    bad = tmp_path / "bad.snap"
    bad.write_bytes(b"NOTASNAP" + b"\0" * 64)
    assert lib.mpm_snapshot_info(str(bad).encode(), None, None, None) == _lib.MPM_E_INVALID
    assert lib.mpm_snapshot_info(str(tmp_path / "missing").encode(), None, None, None) == _lib.MPM_E_INVALID
    short = tmp_path / "short.snap"
    short.write_bytes(b"MPMSNAP1" + struct.pack("<II", 1, ctypes.sizeof(Params)))
    assert lib.mpm_snapshot_info(str(short).encode(), None, None, None) == _lib.MPM_E_INVALID


def _decode_png(data: bytes):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, {}
    idat = b""
    while pos < len(data):
        (length,) = struct.unpack(">I", data[pos:pos + 4])
        ctype = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + length]
        (crc,) = struct.unpack(">I", data[pos + 8 + length:pos + 12 + length])
        assert crc == zlib.crc32(ctype + body) & 0xFFFFFFFF
        if ctype == b"IDAT":
            idat += body
        chunks[ctype] = body
        pos += 12 + length
    w, h, depth, ctype_, _, _, _ = struct.unpack(">IIBBBBB", chunks[b"IHDR"])
    assert depth == 8 and ctype_ == 2 and b"IEND" in chunks
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(h, w, 3)


def test_png_writer(tmp_path):
    from mpm_snow_sim2d_amd.solver import write_png
    rng = np.random.default_rng(1)
    for w, h in ((1, 1), (37, 5), (300, 250)):     # 300x250 needs several stored blocks
        img = rng.integers(0, 256, size=(h, w, 3), dtype=np.uint8)
        path = tmp_path / f"i{w}x{h}.png"
        write_png(str(path), img)
        assert np.array_equal(_decode_png(path.read_bytes()), img)
