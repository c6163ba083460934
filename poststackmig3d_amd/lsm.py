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
iterations; `resume=path` continues from such a file.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

from . import _lib
from .kernel import SurveyParams, demigrate_volume, dot_f64, migrate_volume


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


def cgls(data, vel, params: SurveyParams, iterations: int = 10, model0=None,
         ctx: _lib.Context | None = None, checkpoint: str | None = None,
         checkpoint_every: int = 1, resume: str | None = None,
         stop_after: int | None = None) -> CGResult:
    """CGLS least-squares migration.  data, vel: device tensors
    [n_il][n_xl][n_t].  stop_after: return after that many iterations of this
    call (the state is checkpointed first), for staged runs."""
    import torch

    ctx = ctx or _lib.context()
    F = lambda m: demigrate_volume(m, vel, params, ctx=ctx)       # noqa: E731
    Ft = lambda r: migrate_volume(r, vel, params, ctx=ctx)         # noqa: E731
    dot = lambda a, b: dot_f64(a, b, ctx=ctx)                      # noqa: E731

    if resume is not None:
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

    done_here = 0
    k = k0
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
        if checkpoint is not None and (k % checkpoint_every == 0 or last):
            _save(checkpoint, {"m": m, "r": r, "p": p, "gamma": gamma, "iteration": k,
                               "residual_norms": rn, "gradient_norms": gn})
        if stop_after is not None and done_here >= stop_after:
            break
    return CGResult(model=m, residual_norms=rn, gradient_norms=gn, iterations=k)
