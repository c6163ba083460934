"""Least-squares Kirchhoff migration (CGLS) on the GPU kernels, with
checkpoint / resume across iterations.

Modelling operator F = demigration (psm_demigrate_volume, image -> data), its
adjoint F^T = migration (psm_migrate_volume, data -> image).  CGLS minimises
||F m - d||^2:

    r = d - F m,  s = F^T r,  p = s,  gamma = <s, s>
    repeat:  q = F p;  a = gamma / <q, q>;  m += a p;  r -= a q
             s = F^T r;  gamma' = <s, s>;  p = s + (gamma'/gamma) p

Vectors live in HBM as float32 torch tensors (device memory plumbing); the
inner products and residual norms are accumulated in float64 on the GPU
(psm_dot_f64).  With `checkpoint=path` the full iteration state (m, r, p,
gamma, iteration, residual history) is written every `checkpoint_every`
iterations; `resume=path` continues from such a file.  Checkpoints are
written behind the solver: the state is copied into page-locked host buffers
in stream order and a background thread writes the file (to path.tmp, then
renamed over path, so the file is always a whole checkpoint).  The last
checkpoint of a call may still be in flight when cgls returns; it is waited
for by the next checkpoint written to the same path, by cgls(resume=path),
by wait_checkpoints() and at interpreter exit.
"""
from __future__ import annotations

import atexit
import dataclasses
import math
import os
import threading
from dataclasses import dataclass, field

from . import _lib
from .kernel import SurveyParams, demigrate_volume, dot_f64, migrate_volume, survey_plan


@dataclass
class CGResult:
    model: object                     # torch tensor, image shape
    residual_norms: list = field(default_factory=list)   # ||r_k||, k = 0..iterations
    gradient_norms: list = field(default_factory=list)   # ||F^T r_k||
    iterations: int = 0


def _save(path, state):
    import torch
    tmp = path + ".tmp"
    torch.save(state, tmp)
    os.replace(tmp, path)


# page-locked host buffers kept across calls, keyed by (shape, dtype)
_pinned: dict = {}
_pinned_lock = threading.Lock()

# writers whose last checkpoint was still being written when their cgls call
# returned, by path
_pending: dict = {}
_pending_lock = threading.Lock()


def _take_pending(path):
    with _pending_lock:
        return _pending.pop(path, None)


def wait_checkpoints(path: str | None = None) -> None:
    """Blocks until the checkpoint write still in flight for `path` (every
    path if None) is on disk; re-raises its error, if any."""
    if path is not None:
        writers = [_take_pending(path)]
    else:
        with _pending_lock:
            writers = list(_pending.values())
            _pending.clear()
    for w in writers:
        if w is not None:
            w.close()


atexit.register(wait_checkpoints)


class _CheckpointWriter:
    """Writes checkpoints without stalling the solver.  submit() copies the
    device tensors of the state into page-locked host buffers on torch's
    current stream -- in stream order, so the solver's later in-place updates
    cannot overtake the copy -- and a background thread waits for the copy
    and writes the file (to path.tmp, then renamed over path).  One write is
    in flight at a time; wait() blocks until it is on disk and re-raises its
    error, if any.  A write left in flight by an earlier call on the same
    path is waited for before this writer's first write."""

    def __init__(self, path):
        self.path = path
        self.thread = None
        self.error = None
        self.bufs = {}
        self.prev = _take_pending(path)

    def _finish_prev(self):
        if self.prev is not None:
            prev, self.prev = self.prev, None
            prev.close()

    def _buffer(self, name, t):
        import torch
        b = self.bufs.get(name)
        if b is None or b.shape != t.shape or b.dtype != t.dtype:
            key = (tuple(t.shape), t.dtype)
            with _pinned_lock:
                pool = _pinned.setdefault(key, [])
                b = pool.pop() if pool else None
            if b is None:
                b = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
            self.bufs[name] = b
        return b

    def submit(self, state):
        import torch
        self.wait()
        self._finish_prev()
        host = {}
        on_device = False
        for k, v in state.items():
            if torch.is_tensor(v) and v.is_cuda:
                b = self._buffer(k, v)
                b.copy_(v, non_blocking=True)
                host[k] = b
                on_device = True
            elif torch.is_tensor(v):
                host[k] = v.clone()   # (host tensors: a copy the solver cannot overtake)
            elif isinstance(v, list):
                host[k] = list(v)
            else:
                host[k] = v
        done = None
        if on_device:
            done = torch.cuda.Event()
            done.record()

        def write():
            try:
                if done is not None:
                    done.synchronize()
                _save(self.path, host)
            except BaseException as e:  # re-raised by wait()
                self.error = e

        self.thread = threading.Thread(target=write, name="cgls-checkpoint", daemon=True)
        self.thread.start()

    def wait(self):
        if self.thread is not None:
            self.thread.join()
            self.thread = None
        if self.error is not None:
            e, self.error = self.error, None
            raise e

    def close(self):
        """Waits for the write in flight (and an earlier call's on the same
        path) and returns the page-locked buffers to the pool."""
        try:
            self._finish_prev()
        finally:
            try:
                self.wait()
            finally:
                with _pinned_lock:
                    for b in self.bufs.values():
                        _pinned.setdefault((tuple(b.shape), b.dtype), []).append(b)
                self.bufs = {}

    def detach(self):
        """End of a call: a write still in flight is left to finish behind
        the caller (registered by path, its buffers held until it is waited
        for); otherwise the buffers go back to the pool now."""
        if self.thread is None:
            self.close()
            return
        with _pending_lock:
            earlier = _pending.pop(self.path, None)
            _pending[self.path] = self
        if earlier is not None:   # (a path shared by concurrent calls)
            earlier.close()


def cgls(data, vel, params: SurveyParams, iterations: int = 10, model0=None,
         ctx: _lib.Context | None = None, checkpoint: str | None = None,
         checkpoint_every: int = 1, resume: str | None = None,
         stop_after: int | None = None) -> CGResult:
    """CGLS least-squares migration.  data, vel: device tensors
    [n_il][n_xl][n_t].  stop_after: return after that many iterations of this
    call (the state is checkpointed first), for staged runs.  The last
    checkpoint may still be being written when this returns (see
    wait_checkpoints)."""
    import torch

    ctx = ctx or _lib.context()
    if not params.max_velocity > 0:
        # every application shares vel: probe its maximum once (as each call
        # would, main.cpp:22-64), so the calls skip their synchronous probes
        # and queue behind one another
        params = dataclasses.replace(params, max_velocity=survey_plan(params, vel).max_velocity)
    F = lambda m: demigrate_volume(m, vel, params, ctx=ctx)       # noqa: E731
    Ft = lambda r: migrate_volume(r, vel, params, ctx=ctx)         # noqa: E731
    dot = lambda a, b: dot_f64(a, b, ctx=ctx)                      # noqa: E731

    if resume is not None:
        wait_checkpoints(resume)
        st = torch.load(resume, map_location=data.device, weights_only=True)
        m, r, p = st["m"], st["r"], st["p"]
        gamma = float(st["gamma"])
        k0 = int(st["iteration"])
        rn = list(st["residual_norms"])
        gn = list(st["gradient_norms"])
    else:
        if model0 is None:
            m = None
            r = data.clone()
        else:
            m = model0.clone()
            r = data - F(m)
        s = Ft(r)
        if m is None:
            m = torch.zeros_like(s)
        p = s.clone()
        gamma = dot(s, s)
        k0 = 0
        rn = [math.sqrt(dot(r, r))]
        gn = [math.sqrt(gamma)]

    writer = _CheckpointWriter(checkpoint) if checkpoint is not None else None
    try:
        k = _iterate(F, Ft, dot, m, r, p, gamma, k0, rn, gn, iterations, writer,
                     checkpoint_every, stop_after)
    except BaseException:
        if writer is not None:
            try:
                writer.close()
            except Exception:
                pass   # (the solver's error is the one raised)
        raise
    if writer is not None:
        writer.detach()   # the last write finishes behind the caller
    return CGResult(model=m, residual_norms=rn, gradient_norms=gn, iterations=k)


def _iterate(F, Ft, dot, m, r, p, gamma, k, rn, gn, iterations, writer, checkpoint_every,
             stop_after):
    """CGLS iterations k .. iterations - 1 (m, r, p updated in place, rn and gn
    appended); returns the iteration count reached."""
    done_here = 0
    while k < iterations:
        q = F(p)
        qq = dot(q, q)
        if qq == 0.0 or gamma == 0.0:
            break
        a = gamma / qq
        m.add_(p, alpha=a)
        r.add_(q, alpha=-a)
        s = Ft(r)
        g_new = dot(s, s)
        p.mul_(g_new / gamma).add_(s)
        gamma = g_new
        k += 1
        done_here += 1
        rn.append(math.sqrt(dot(r, r)))
        gn.append(math.sqrt(gamma))
        last = k == iterations or (stop_after is not None and done_here >= stop_after)
        if writer is not None and (k % checkpoint_every == 0 or last):
            writer.submit({"m": m, "r": r, "p": p, "gamma": gamma, "iteration": k,
                           "residual_norms": rn, "gradient_norms": gn})
        if stop_after is not None and done_here >= stop_after:
            break
    return k
