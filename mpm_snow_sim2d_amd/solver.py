"""Host-side mirror of the reference's solver interface over the C-ABI.

The reference selects ``GPUSolver`` or ``CPUSolver`` at compile time and uses them
through ``Solver()``, ``Step()`` and ``GetParticlePositions()`` (src/main.cpp:22-29,
src/gpu/GPUSolver.h:10-17, src/cpu/CPUSolver.h:10-27).  ``GPUSolver`` here keeps
those names and meanings on top of libmpm_hip.so, and adds the stage-level calls
(``p2g``, ``grid_update``, ``g2p``, ``reset``), state and grid readback, binning and
timing.  Errors raise ``MPMError`` (the reference throws from its constructor,
GPUSolver.cpp:39-50).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import CParticles, MPMError, check
from .params import COptions, Options, Params, reference_params
from .scenes import ParticleSet, reference_snowballs


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f32(a, shape) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, np.float32))
    if a.size != int(np.prod(shape)):
        raise ValueError(f"expected {shape}, got {a.shape}")
    return a.reshape(shape)


class GPUSolver:
    """MLS-MPM snow solver on the GPU (drop-in for the reference's GPUSolver)."""

    def __init__(self, params: Params | None = None, particles: ParticleSet | None = None,
                 options: Options | None = None, C=None, F=None, Jp=None):
        lib = _lib.load()
        self.params = params if params is not None else reference_params(2)
        if particles is None:
            particles = reference_snowballs(self.params.dim)
        d = self.params.dim
        if particles.dim != d:
            raise ValueError("particle dimension does not match params.dim")
        n = particles.n
        self.n = n
        self.dim = d
        self._keep = [_f32(particles.x, (d, n)), _f32(particles.v, (d, n)),
                      _f32(particles.mass, (n,)), _f32(particles.vol, (n,))]
        extra = [None if C is None else _f32(C, (d * d, n)),
                 None if F is None else _f32(F, (d * d, n)),
                 None if Jp is None else _f32(Jp, (n,))]
        cp = CParticles(n, _ptr(self._keep[0]), _ptr(self._keep[1]), _ptr(self._keep[2]),
                        _ptr(self._keep[3]), _ptr(extra[0]), _ptr(extra[1]), _ptr(extra[2]))
        opts = COptions.of(options or Options())
        h = ctypes.c_void_p()
        check(lib.mpm_create(ctypes.byref(self.params), ctypes.byref(cp), ctypes.byref(opts), ctypes.byref(h)))
        self._h = h
        self._lib = lib
        self._cached_positions = np.zeros((n, d), np.float32)

    # -- lifetime
    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.mpm_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - interpreter teardown order
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- the reference's interface
    def Step(self, n_substeps: int = 1) -> None:
        """GPUSolver::Step (GPUSolver.cpp:63-80): one (or n) substeps, asynchronous."""
        check(self._lib.mpm_step(self._h, int(n_substeps)))

    def GetParticlePositions(self) -> np.ndarray:
        """GPUSolver::GetParticlePositions (GPUSolver.cpp:276-295): synchronises and
        returns an (n, dim) array owned by the solver, valid until the next call."""
        check(self._lib.mpm_get_positions(self._h, _ptr(self._cached_positions)))
        return self._cached_positions

    step = Step

    # -- stages (one per dispatch of GPUSolver::Step)
    def p2g(self) -> None:
        check(self._lib.mpm_p2g(self._h))

    def grid_update(self) -> None:
        check(self._lib.mpm_grid_update(self._h))

    def g2p(self) -> None:
        check(self._lib.mpm_g2p(self._h))

    def reset(self) -> None:
        check(self._lib.mpm_reset_grid(self._h))

    def sort(self) -> None:
        check(self._lib.mpm_sort(self._h))

    def sync(self) -> None:
        check(self._lib.mpm_sync(self._h))

    def set_graph(self, on: bool = True) -> None:
        """Replay whole binning periods of ``Step`` from one recorded hipGraph."""
        check(self._lib.mpm_set_graph(self._h, int(on)))

    # -- state
    def get_state(self) -> dict:
        d, n = self.dim, self.n
        x = np.empty((d, n), np.float32)
        v = np.empty((d, n), np.float32)
        C = np.empty((d * d, n), np.float32)
        F = np.empty((d * d, n), np.float32)
        Jp = np.empty(n, np.float32)
        check(self._lib.mpm_get_state(self._h, _ptr(x), _ptr(v), _ptr(C), _ptr(F), _ptr(Jp)))
        return {"x": x, "v": v, "C": C, "F": F, "Jp": Jp}

    def set_state(self, x=None, v=None, C=None, F=None, Jp=None) -> None:
        d, n = self.dim, self.n
        arrs = [None if x is None else _f32(x, (d, n)), None if v is None else _f32(v, (d, n)),
                None if C is None else _f32(C, (d * d, n)), None if F is None else _f32(F, (d * d, n)),
                None if Jp is None else _f32(Jp, (n,))]
        check(self._lib.mpm_set_state(self._h, *[_ptr(a) for a in arrs]))

    def get_grid(self) -> dict:
        G = self.params.num_nodes
        d = self.dim
        m = np.empty(G, np.float32)
        mom = np.empty((d, G), np.float32)
        vc = np.empty((d, G), np.float32)
        vf = np.empty((d, G), np.float32)
        check(self._lib.mpm_get_grid(self._h, _ptr(m), _ptr(mom), _ptr(vc), _ptr(vf)))
        return {"mass": m, "mom": mom, "vcol": vc, "vfric": vf}

    def get_grid_force(self) -> np.ndarray:
        """CPU-path scheme: the explicit grid force after p2g(), (dim, nodes)."""
        f = np.empty((self.dim, self.params.num_nodes), np.float32)
        check(self._lib.mpm_get_grid_force(self._h, _ptr(f)))
        return f

    def get_sort(self) -> tuple[np.ndarray, np.ndarray]:
        keys = np.empty(self.n, np.uint32)
        perm = np.empty(self.n, np.uint32)
        check(self._lib.mpm_get_sort(self._h, _ptr(keys), _ptr(perm)))
        return keys, perm

    def check_finite(self) -> int:
        bad = ctypes.c_int64(0)
        code = self._lib.mpm_check_finite(self._h, ctypes.byref(bad))
        if code not in (_lib.MPM_OK, _lib.MPM_E_NONFINITE):
            check(code)
        return int(bad.value)

    # -- timing
    def set_timing(self, on: bool = True, kernels=None) -> None:
        """Per-kernel HIP-event timing; kernels: an iterable of KERNEL_* ids to time
        only those (fewer event records between launches)."""
        if kernels is None:
            check(self._lib.mpm_set_timing(self._h, int(on)))
        else:
            mask = 0
            for k in kernels:
                mask |= 1 << int(k)
            check(self._lib.mpm_set_timing_mask(self._h, mask if on else 0))

    def set_timing_every(self, every: int) -> None:
        """Time only every k-th launch of each timed kernel (mpm_set_timing_every)."""
        check(self._lib.mpm_set_timing_every(self._h, int(every)))

    def reset_timers(self) -> None:
        check(self._lib.mpm_reset_timers(self._h))

    def kernel_time(self, kernel: int) -> tuple[float, int]:
        ms = ctypes.c_double(0.0)
        k = ctypes.c_int64(0)
        check(self._lib.mpm_get_kernel_time(self._h, int(kernel), ctypes.byref(ms), ctypes.byref(k)))
        return float(ms.value), int(k.value)

    # -- snapshots and frames
    def save_snapshot(self, path: str) -> None:
        """mpm_save_snapshot: parameters, substep count and state (snapshot.py reads it)."""
        check(self._lib.mpm_save_snapshot(self._h, str(path).encode()))

    @classmethod
    def from_snapshot(cls, path: str, options: Options | None = None) -> "GPUSolver":
        lib = _lib.load()
        h = ctypes.c_void_p()
        opts = COptions.of(options or Options())
        check(lib.mpm_load_snapshot(str(path).encode(), ctypes.byref(opts), ctypes.byref(h)))
        self = cls.__new__(cls)
        self._h, self._lib = h, lib
        P = Params()
        n = ctypes.c_int64(0)
        check(lib.mpm_snapshot_info(str(path).encode(), ctypes.byref(P), ctypes.byref(n), None))
        self.params, self.n, self.dim = P, int(n.value), P.dim
        self._keep = []
        self._cached_positions = np.zeros((self.n, self.dim), np.float32)
        return self

    def render_frame(self, width: int = 1200, height: int = 600, camera: dict | None = None,
                     point_size: int = 1) -> np.ndarray:
        """mpm_render_frame: (height, width, 3) uint8, top row first (Visualize +
        saveFrame, src/main.cpp:56-160, 177-205)."""
        from ._lib import CCamera
        cam = CCamera()
        check(self._lib.mpm_camera_default(ctypes.byref(cam)))
        for k, v in (camera or {}).items():
            setattr(cam, k, float(v))
        img = np.empty((height, width, 3), np.uint8)
        check(self._lib.mpm_render_frame(self._h, ctypes.byref(cam), int(width), int(height), int(point_size),
                                         _ptr(img)))
        return img


def write_png(path: str, rgb: np.ndarray) -> None:
    """mpm_write_png: an (h, w, 3) uint8 image, top row first."""
    lib = _lib.load()
    img = np.ascontiguousarray(rgb, np.uint8)
    check(lib.mpm_write_png(str(path).encode(), img.shape[1], img.shape[0], _ptr(img)))


def save_frames(solver: GPUSolver, out_dir: str, frames: int, substeps_per_frame: int, **render) -> list[str]:
    """The reference's main loop (src/main.cpp:272-296) without a window: advance,
    render, save frame_<k>.png.  Returns the written paths."""
    import os
    os.makedirs(out_dir, exist_ok=True)
    paths = []
    for k in range(frames):
        solver.Step(substeps_per_frame)
        path = os.path.join(out_dir, f"frame_{k}.png")
        write_png(path, solver.render_frame(**render))
        paths.append(path)
    return paths


def _c_particles(ps: ParticleSet, keep: list) -> CParticles:
    d, n = ps.dim, ps.n
    arrs = [_f32(ps.x, (d, n)), _f32(ps.v, (d, n)), _f32(ps.mass, (n,)), _f32(ps.vol, (n,))]
    keep.extend(arrs)
    return CParticles(n, *[_ptr(a) for a in arrs], None, None, None)


def _scene_arrays(scenes):
    if not scenes:
        raise ValueError("a batch needs at least one scene")
    keep: list = []
    P = (Params * len(scenes))(*[p for p, _ in scenes])
    C = (CParticles * len(scenes))(*[_c_particles(ps, keep) for _, ps in scenes])
    return P, C, keep


class BatchSolver:
    """Many independent scenes advanced together (``mpm_batch_create``).

    ``scenes`` is a list of (Params, ParticleSet); state arrays hold the scenes one
    after another in that order (``scene_range``)."""

    def __init__(self, scenes, options: Options | None = None, graph: bool = False):
        lib = _lib.load()
        self._lib = lib
        self.scenes = list(scenes)
        self.dim = self.scenes[0][0].dim
        P, C, keep = _scene_arrays(self.scenes)
        opts = COptions.of(options or Options())
        h = ctypes.c_void_p()
        check(lib.mpm_batch_create(len(self.scenes), P, C, ctypes.byref(opts), ctypes.byref(h)))
        self._h = h
        del keep
        n = ctypes.c_int64(0)
        check(lib.mpm_num_particles(self._h, ctypes.byref(n)))
        self.n = int(n.value)
        if graph:
            self.set_graph(True)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.mpm_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - interpreter teardown order
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def num_scenes(self) -> int:
        k = ctypes.c_int(0)
        check(self._lib.mpm_num_scenes(self._h, ctypes.byref(k)))
        return int(k.value)

    def scene_range(self, scene: int) -> tuple[int, int]:
        first = ctypes.c_int64(0)
        count = ctypes.c_int64(0)
        check(self._lib.mpm_scene_range(self._h, int(scene), ctypes.byref(first), ctypes.byref(count)))
        return int(first.value), int(count.value)

    def Step(self, n_substeps: int = 1) -> None:
        check(self._lib.mpm_step(self._h, int(n_substeps)))

    step = Step

    def set_graph(self, on: bool = True) -> None:
        check(self._lib.mpm_set_graph(self._h, int(on)))

    def sync(self) -> None:
        check(self._lib.mpm_sync(self._h))

    def get_state(self) -> dict:
        d, n = self.dim, self.n
        x = np.empty((d, n), np.float32)
        v = np.empty((d, n), np.float32)
        C = np.empty((d * d, n), np.float32)
        F = np.empty((d * d, n), np.float32)
        Jp = np.empty(n, np.float32)
        check(self._lib.mpm_get_state(self._h, _ptr(x), _ptr(v), _ptr(C), _ptr(F), _ptr(Jp)))
        return {"x": x, "v": v, "C": C, "F": F, "Jp": Jp}

    def scene_state(self, state: dict, scene: int) -> dict:
        a, c = self.scene_range(scene)
        return {k: v[..., a:a + c] for k, v in state.items()}

    def set_state(self, x=None, v=None, C=None, F=None, Jp=None) -> None:
        d, n = self.dim, self.n
        arrs = [None if x is None else _f32(x, (d, n)), None if v is None else _f32(v, (d, n)),
                None if C is None else _f32(C, (d * d, n)), None if F is None else _f32(F, (d * d, n)),
                None if Jp is None else _f32(Jp, (n,))]
        check(self._lib.mpm_set_state(self._h, *[_ptr(a) for a in arrs]))

    def set_timing(self, on: bool = True, kernels=None) -> None:
        """Per-kernel HIP-event timing; kernels: an iterable of KERNEL_* ids to time
        only those (fewer event records between launches)."""
        if kernels is None:
            check(self._lib.mpm_set_timing(self._h, int(on)))
        else:
            mask = 0
            for k in kernels:
                mask |= 1 << int(k)
            check(self._lib.mpm_set_timing_mask(self._h, mask if on else 0))

    def set_timing_every(self, every: int) -> None:
        """Time only every k-th launch of each timed kernel (mpm_set_timing_every)."""
        check(self._lib.mpm_set_timing_every(self._h, int(every)))

    def reset_timers(self) -> None:
        check(self._lib.mpm_reset_timers(self._h))

    def kernel_time(self, kernel: int) -> tuple[float, int]:
        ms = ctypes.c_double(0.0)
        k = ctypes.c_int64(0)
        check(self._lib.mpm_get_kernel_time(self._h, int(kernel), ctypes.byref(ms), ctypes.byref(k)))
        return float(ms.value), int(k.value)

    def check_finite(self) -> int:
        bad = ctypes.c_int64(0)
        code = self._lib.mpm_check_finite(self._h, ctypes.byref(bad))
        if code not in (_lib.MPM_OK, _lib.MPM_E_NONFINITE):
            check(code)
        return int(bad.value)


def autotune(scenes, trial_substeps: int = 64, graph: bool = True) -> tuple[Options, float]:
    """``mpm_autotune``: the fastest (tile_particles, sort_every) for these scenes and
    its time per substep in ms.  ``scenes`` is a list of (Params, ParticleSet)."""
    lib = _lib.load()
    P, C, keep = _scene_arrays(list(scenes))
    best = COptions()
    ms = ctypes.c_double(0.0)
    check(lib.mpm_autotune(len(scenes), P, C, int(trial_substeps), int(graph), ctypes.byref(best),
                           ctypes.byref(ms)))
    del keep
    return Options(best.sort_every, best.fused, best.tile_particles), float(ms.value)


def sort_pairs(keys: np.ndarray, key_bits: int = 32) -> tuple[np.ndarray, np.ndarray]:
    """Stable device sort of uint32 keys; returns (sorted keys, permutation)."""
    lib = _lib.load()
    k = np.ascontiguousarray(keys, np.uint32)
    n = k.size
    ok = np.empty(n, np.uint32)
    op = np.empty(n, np.uint32)
    check(lib.mpm_sort_pairs_host(_ptr(k), n, int(key_bits), _ptr(ok), _ptr(op)))
    return ok, op


__all__ = ["BatchSolver", "GPUSolver", "MPMError", "autotune", "save_frames", "sort_pairs", "write_png"]
