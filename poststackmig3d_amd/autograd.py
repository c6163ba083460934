"""torch.autograd wrapper of the migration operator (test and LSM harness).

    image = migrate(data, alpha, vel, params)

forward  : psm_migrate_volume at velocity scale alpha (mask at alpha_mask)
backward : d data  <- psm_demigrate_volume(grad)               (M^T g)
           d alpha <- <psm_migrate_volume_dalpha(data), grad>  (float64 dot)

so torch.autograd.grad / gradcheck-style tests exercise the hand-written
adjoint and tangent kernels through the C ABI.  The aperture mask is held at
alpha_mask (default 1): the operator is then smooth in alpha between the
integer crossings of the sample index.
"""
from __future__ import annotations

import torch

from . import _lib
from .kernel import (SurveyParams, demigrate_volume, dot_f64, migrate_volume,
                     migrate_volume_dalpha)


class _Migrate(torch.autograd.Function):
    @staticmethod
    def forward(fctx, data, alpha, vel, params, ctx, alpha_mask):
        a = float(alpha)
        ctx.set_velocity_scale(a, alpha_mask)
        try:
            out = migrate_volume(data.contiguous(), vel, params, ctx=ctx)
        finally:
            ctx.set_velocity_scale(1.0, 1.0)
        fctx.save_for_backward(data)
        fctx.vel, fctx.params, fctx.pctx, fctx.alpha, fctx.alpha_mask = vel, params, ctx, a, alpha_mask
        fctx.alpha_meta = (alpha.dtype, alpha.device) if torch.is_tensor(alpha) else (torch.float64, None)
        return out

    @staticmethod
    def backward(fctx, grad):
        (data,) = fctx.saved_tensors
        ctx = fctx.pctx
        g = grad.contiguous()
        ctx.set_velocity_scale(fctx.alpha, fctx.alpha_mask)
        try:
            gd = demigrate_volume(g, fctx.vel, fctx.params, ctx=ctx) \
                if fctx.needs_input_grad[0] else None
            ga = None
            if fctx.needs_input_grad[1]:
                tan = migrate_volume_dalpha(data.contiguous(), fctx.vel, fctx.params, ctx=ctx)
                dtype, dev = fctx.alpha_meta
                ga = torch.tensor(dot_f64(tan, g, ctx=ctx), dtype=dtype, device=dev)
        finally:
            ctx.set_velocity_scale(1.0, 1.0)
        return gd, ga, None, None, None, None


def migrate(data, alpha, vel, params: SurveyParams, ctx: _lib.Context | None = None,
            alpha_mask: float = 1.0):
    """Differentiable migration of a float32 CUDA tensor `data` at velocity
    scale `alpha` (a 0-d tensor to differentiate with respect to)."""
    ctx = ctx or _lib.context()
    if not torch.is_tensor(alpha):
        alpha = torch.tensor(float(alpha), dtype=torch.float64)
    return _Migrate.apply(data, alpha, vel, params, ctx, float(alpha_mask))
