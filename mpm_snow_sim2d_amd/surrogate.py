"""Learned grid surrogate over the C-ABI (``mpm_surrogate_*``): a BF16 U-Net on the
matrix cores that predicts the grid velocity and displacement k substeps ahead
(transfers or least-squares fits of the particles' values; the displacement,
when the macro step uses it, as a second image of the batch told apart by a
constant mode plane), trained on the
device from trajectories the solver generates, and a hybrid rollout that uses it
where its residual test passes and the plain solver elsewhere.

``Surrogate`` wraps one network with its device-resident training set;
``generate`` fills the set from a batch of scenes; ``fit`` runs Adam steps;
``rollout`` advances a single-scene solver in macro steps of k substeps.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from . import _lib
from ._lib import check

ADVANCE_PLAIN = _lib.ADVANCE_PLAIN
ADVANCE_NETWORK = _lib.ADVANCE_NETWORK
ADVANCE_BALLISTIC = _lib.ADVANCE_BALLISTIC
CONV_AUTO, CONV_STREAM, CONV_LDS = _lib.CONV_AUTO, _lib.CONV_STREAM, _lib.CONV_LDS
MACRO_X, MACRO_F = 1, 2   # MPM_MACRO_DISPLACEMENT_X, MPM_MACRO_DISPLACEMENT_F


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Surrogate:
    """One U-Net (height x width grid nodes) and `capacity` training slots."""

    def __init__(self, height: int, width: int, max_batch: int = 8, capacity: int = 0, seed: int = 0,
                 d_ref: float = 0.1):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        check(self._lib.mpm_surrogate_create(int(height), int(width), int(max_batch), int(capacity),
                                             ctypes.c_uint64(seed), ctypes.byref(h)))
        self._h = h
        self.h, self.w, self.max_batch, self.capacity = height, width, max_batch, capacity
        self.set_scale(d_ref)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.mpm_surrogate_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- parameters
    @property
    def num_params(self) -> int:
        n = ctypes.c_int64(0)
        check(self._lib.mpm_surrogate_num_params(self._h, ctypes.byref(n)))
        return int(n.value)

    def get_params(self) -> np.ndarray:
        p = np.empty(self.num_params, np.float32)
        check(self._lib.mpm_surrogate_get_params(self._h, _ptr(p)))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.num_params:
            raise ValueError("parameter vector has the wrong size")
        check(self._lib.mpm_surrogate_set_params(self._h, _ptr(p)))

    def set_scale(self, d_ref: float) -> None:
        self.d_ref = float(d_ref)
        check(self._lib.mpm_surrogate_set_scale(self._h, ctypes.c_float(d_ref)))

    def set_macro(self, flags: int) -> None:
        """Fields of the macro step: MACRO_X moves particles by the predicted
        displacement (else by the trapezoid of the velocities), MACRO_F advances F_E by
        its gradient (else by T C'); velocity and C always from the predicted velocity."""
        check(self._lib.mpm_surrogate_set_macro(self._h, int(flags)))

    def set_targets(self, refine_iters: int) -> None:
        """Jacobi steps of the least-squares fit applied to later target captures
        (0: the plain transfers)."""
        check(self._lib.mpm_surrogate_set_targets(self._h, int(refine_iters)))

    def set_conv(self, kernel: int) -> None:
        """Forward convolution kernel: CONV_AUTO (by layer shape), CONV_STREAM
        (register-streaming) or CONV_LDS (LDS-tiled)."""
        check(self._lib.mpm_surrogate_set_conv(self._h, int(kernel)))

    # -- the network on host arrays
    def forward(self, input_bits: np.ndarray, wm: np.ndarray) -> np.ndarray:
        nb = input_bits.shape[0]
        x = np.ascontiguousarray(input_bits, np.uint16)
        w = np.ascontiguousarray(wm, np.float32)
        out = np.empty((nb, self.h * self.w, 2), np.float32)
        check(self._lib.mpm_surrogate_forward(self._h, nb, _ptr(x), _ptr(w), _ptr(out)))
        return out

    def gradient(self, input_bits: np.ndarray, wm: np.ndarray, target: np.ndarray, grad: bool = True):
        nb = input_bits.shape[0]
        x = np.ascontiguousarray(input_bits, np.uint16)
        w = np.ascontiguousarray(wm, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        loss = np.zeros(1, np.float32)
        g = np.empty(self.num_params, np.float32) if grad else None
        check(self._lib.mpm_surrogate_gradient(self._h, nb, _ptr(x), _ptr(w), _ptr(t), _ptr(loss), _ptr(g)))
        return float(loss[0]), g

    # -- training set
    def record(self, solver, scene: int, slot: int, what: int, k: int) -> None:
        check(self._lib.mpm_surrogate_record(self._h, solver._h, int(scene), int(slot), int(what), int(k)))

    def get_sample(self, slot: int):
        data = np.empty((9, self.h * self.w), np.float32)
        sc = np.empty(6, np.float32)
        check(self._lib.mpm_surrogate_get_sample(self._h, int(slot), _ptr(data), _ptr(sc)))
        keys = ("particle_mass", "grid_mass", "band_mass", "T", "gravity", "m_ref")
        return data, dict(zip(keys, map(float, sc)))

    def get_scalars(self, slot: int) -> dict:
        """The scalars of a slot only (no copy of its images)."""
        sc = np.empty(6, np.float32)
        check(self._lib.mpm_surrogate_get_sample(self._h, int(slot), None, _ptr(sc)))
        keys = ("particle_mass", "grid_mass", "band_mass", "T", "gravity", "m_ref")
        return dict(zip(keys, map(float, sc)))

    def train(self, slots, lr: float) -> float:
        s = np.ascontiguousarray(slots, np.int32)
        loss = np.zeros(1, np.float32)
        check(self._lib.mpm_surrogate_train(self._h, _ptr(s), s.size, ctypes.c_float(lr), _ptr(loss)))
        return float(loss[0])

    def loss(self, slots) -> float:
        return self.train(slots, 0.0)

    def baseline_loss(self, slots) -> float:
        s = np.ascontiguousarray(slots, np.int32)
        loss = np.zeros(1, np.float32)
        check(self._lib.mpm_surrogate_baseline_loss(self._h, _ptr(s), s.size, _ptr(loss)))
        return float(loss[0])

    def time(self, nb: int, iters: int = 10, train: bool = False) -> float:
        """Average device ms of a forward (or forward + backward) pass on nb images
        (a macro step runs 1, or 2 with displacement; a training step as many per slot)."""
        ms = ctypes.c_double(0.0)
        check(self._lib.mpm_surrogate_time(self._h, int(nb), int(iters), int(train), ctypes.byref(ms)))
        return float(ms.value)

    def flops(self, nb: int = 1) -> float:
        """Matrix-core FLOPs of one forward pass on nb images (3x3 convolutions as
        specified, 32 input channels at the first layer, plus the 1x1 head)."""
        hw = self.h * self.w
        spec = [(32, 32, 0), (32, 64, 1), (64, 128, 2), (128, 128, 3), (256, 128, 2), (192, 64, 1), (96, 32, 0)]
        f = sum(2.0 * 9 * ci * co * hw / 4 ** lvl for ci, co, lvl in spec) + 2.0 * 32 * 2 * hw
        return f * nb

    # -- rollout
    def advance(self, solver, k: int, mode: int, tol: float = 1e-3) -> tuple[int, float]:
        used = ctypes.c_int(0)
        res = ctypes.c_float(0.0)
        check(self._lib.mpm_surrogate_advance(self._h, solver._h, int(k), int(mode), ctypes.c_float(tol),
                                              ctypes.byref(used), ctypes.byref(res)))
        return int(used.value), float(res.value)

    def rollout_steps(self, solver, k: int, n_macro: int, mode: int, tol: float = 1e-3) -> tuple[int, int, float]:
        """n_macro macro steps as n_macro advance() calls would take them, the ones
        after the solver's first macro step replayed from one recorded graph with the
        residual test on the device.  Returns (macro steps taken, plain fallbacks,
        largest residual); raises MPMError (MPM_E_STATE) when a test fails after a
        macro step (the particles then hold the state before that step)."""
        nn, np_, res = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_float(0.0)
        check(self._lib.mpm_surrogate_rollout(self._h, solver._h, int(k), int(n_macro), int(mode), ctypes.c_float(tol),
                                              ctypes.byref(nn), ctypes.byref(np_), ctypes.byref(res)))
        return int(nn.value), int(np_.value), float(res.value)


def generate(sur: Surrogate, scenes, substeps: int, k: int = 8, first_slot: int = 0, options=None,
             graph: bool = True, wall_tol: float | None = None) -> list[list[int]]:
    """Run the scenes (list of (Params, ParticleSet), all on the surrogate's grid)
    together for `substeps` substeps and record one (input, target) sample per scene
    every k substeps.  Returns the slots of each scene, in time order.  With
    wall_tol, samples whose wall-band mass share exceeds it (where the rollout falls
    back to the plain solver, and the target holds wall impulses the projected
    prediction cannot represent) are left out of the returned lists."""
    from .params import Options
    from .solver import BatchSolver
    nmacro = substeps // k
    need = len(scenes) * nmacro
    if first_slot + need > sur.capacity:
        raise ValueError(f"{need} samples do not fit in the {sur.capacity - first_slot} free slots")
    slots = [[] for _ in scenes]
    with BatchSolver(scenes, options or Options(sort_every=k, tile_particles=256), graph=graph) as b:
        slot = first_slot
        for _ in range(nmacro):
            for i in range(len(scenes)):
                sur.record(b, i, slot + i, 0, k)
            b.Step(k)
            for i in range(len(scenes)):
                sur.record(b, i, slot + i, 1, k)
                if wall_tol is not None:
                    sc = sur.get_scalars(slot + i)
                    pm = max(sc["particle_mass"], 1e-30)
                    if sc["band_mass"] / pm > wall_tol or abs(pm - sc["grid_mass"]) / pm > wall_tol:
                        continue
                slots[i].append(slot + i)
            slot += len(scenes)
    return slots


def fit(sur: Surrogate, train_slots, steps: int, batch: int = 8, lr: float = 1e-3, seed: int = 0,
        lr_final: float | None = None, log_every: int = 0, log=print) -> list[float]:
    """Adam steps on random batches of the training slots (cosine decay of the
    learning rate to lr_final)."""
    rng = np.random.default_rng(seed)
    pool = np.asarray(train_slots, np.int32)
    lr_final = lr * 0.1 if lr_final is None else lr_final
    losses = []
    for it in range(steps):
        lr_t = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * it / max(steps - 1, 1)))
        sel = rng.choice(pool, size=min(batch, pool.size), replace=False)
        want = log_every and (it % log_every == 0 or it == steps - 1)
        if want:
            losses.append(sur.train(sel, lr_t))
            log(f"[fit] step {it} loss {losses[-1]:.5f} lr {lr_t:.2e}")
        else:
            sur.train(sel, lr_t)
    return losses


def train_streaming(sur: Surrogate, scene_fn, chunks: int, substeps: int, k: int = 8, first_slot: int = 0,
                    steps_per_chunk: int = 1000, batch: int = 8, lr: float = 1e-3, lr_final: float | None = None,
                    wall_tol: float | None = 1e-3, seed: int = 0, options=None, eval_fn=None, epochs: int = 1,
                    patience: int = 0, log=print) -> dict:
    """Training on more trajectories than the device holds at once: chunk c takes
    the scenes scene_fn(c) (a list of (Params, ParticleSet)), runs them for
    `substeps` substeps while recording one sample per scene every k substeps into
    the slots from `first_slot` on (overwriting the previous chunk's), then takes
    `steps_per_chunk` Adam steps on random batches of those samples; `epochs` passes
    over the chunks (each pass regenerates them).  The learning rate decays on one
    cosine over all steps.  With eval_fn (a function of the surrogate returning a
    held-out loss) the loss is taken after every chunk, the parameters of the lowest
    are kept and restored at the end, and with patience > 0 training stops once
    that many chunks in a row have not lowered it by 0.5%.  Returns trajectory,
    sample and step counts, the generation / training times and the losses."""
    rng = np.random.default_rng(seed)
    lr_final = lr * 0.1 if lr_final is None else lr_final
    total = max(epochs * chunks * steps_per_chunk - 1, 1)
    it, samples, gen_s, train_s, trajectories = 0, 0, 0.0, 0.0, 0
    evals, best, best_p, stale = [], math.inf, None, 0
    for e in range(epochs * chunks):
        c = e % chunks
        scenes = scene_fn(c)
        t0 = time.perf_counter()
        slots = generate(sur, scenes, substeps, k, first_slot=first_slot, options=options, wall_tol=wall_tol)
        gen_s += time.perf_counter() - t0
        if e < chunks:
            trajectories += len(scenes)
        pool = np.asarray([s for ss in slots for s in ss], np.int32)
        samples += int(pool.size)
        if pool.size == 0:
            continue
        t0 = time.perf_counter()
        last, lr_t = 0.0, lr
        for j in range(steps_per_chunk):
            lr_t = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * it / total))
            sel = rng.choice(pool, size=min(batch, pool.size), replace=False)
            if j == steps_per_chunk - 1:
                last = sur.train(sel, lr_t)
            else:
                sur.train(sel, lr_t)
            it += 1
        sur.loss(pool[:1])                      # waits for the chunk's last step
        train_s += time.perf_counter() - t0
        ev = ""
        if eval_fn is not None:
            evals.append(float(eval_fn(sur)))
            ev = f", held-out {evals[-1]:.5f}"
            if evals[-1] < best * 0.995:
                stale = 0
            else:
                stale += 1
            if evals[-1] < best:
                best, best_p = evals[-1], sur.get_params()
        log(f"[stream] chunk {e + 1}/{epochs * chunks}: {len(scenes)} trajectories, {pool.size} samples, "
            f"{steps_per_chunk} steps, last batch loss {last:.4f}, lr {lr_t:.2e}{ev}")
        if patience and stale >= patience:
            log(f"[stream] held-out loss has not fallen for {patience} chunks: stopping")
            break
    if best_p is not None:
        sur.set_params(best_p)
    return {"trajectories": trajectories, "samples": samples, "steps": it, "generate_s": gen_s, "train_s": train_s,
            "eval": evals, "best": best if best_p is not None else None}


def rollout(sur: Surrogate | None, params, ps, substeps: int, k: int = 8, mode: int = ADVANCE_NETWORK,
            tol: float = 1e-3, options=None, graph: bool = True) -> dict:
    """Advance one scene by `substeps` (a multiple of k) in macro steps; mode
    ADVANCE_PLAIN runs the plain solver only.  With graph, the plain solver replays
    hipGraphs of its substeps and the macro steps after the first replay a graph of
    one macro step (mpm_surrogate_rollout); without, every macro step is one
    advance() call.  Returns the final positions (n, 2), the wall time, and how
    many macro steps each mode took."""
    from .params import Options
    from .solver import GPUSolver
    s = GPUSolver(params, ps, options or Options(sort_every=k, tile_particles=256))
    if graph:
        s.set_graph(True)
    counts = {ADVANCE_PLAIN: 0, ADVANCE_NETWORK: 0, ADVANCE_BALLISTIC: 0}
    stopped = None
    s.sync()
    t0 = time.perf_counter()
    try:
        if mode == ADVANCE_PLAIN or sur is None:
            s.Step(substeps // k * k)
            counts[ADVANCE_PLAIN] = substeps // k
        elif graph:
            nn, np_, _ = sur.rollout_steps(s, k, substeps // k, mode, tol)
            counts[mode] += nn
            counts[ADVANCE_PLAIN] += np_
        else:
            for _ in range(substeps // k):
                used, _ = sur.advance(s, k, mode, tol)
                counts[used] += 1
    except _lib.MPMError as e:
        # a residual test failed after a macro step: the rollout stops there
        if e.code != _lib.MPM_E_STATE:
            s.close()
            raise
        stopped = str(e)
    s.sync()
    wall = time.perf_counter() - t0
    x = s.GetParticlePositions().copy()
    bad = s.check_finite()
    s.close()
    return {"x": x, "wall": wall, "counts": counts, "nonfinite": bad, "stopped": stopped}
