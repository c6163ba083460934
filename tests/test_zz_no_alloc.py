"""psm_migrate_volume_noalloc: the same migration as psm_migrate_volume on
device memory, with its workspace in the library's static array instead of
allocated buffers.  Called through the C ABI; two calls in a row, the second
on a smaller survey, each bit-equal to psm_migrate_volume's result (the two
entry points run the same kernels on the same inputs), and a survey larger
than the workspace rejected with PSM_ERR_ARG before any work is queued."""
import ctypes as C

import numpy as np
import pytest

import poststackmig3d_amd as P
from poststackmig3d_amd import synth

pytestmark = pytest.mark.gpu

PSM_ERR_ARG = 1


def _device(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to("cuda:0")


def _noalloc(ctx, prm, data, vel, out):
    s = prm.to_c()
    return ctx.lib.psm_migrate_volume_noalloc(ctx.h, C.byref(s), data.data_ptr(), vel.data_ptr(),
                                              out.data_ptr())


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_noalloc_equals_migrate_volume(storage):
    import torch
    ctx = P.context(0)
    ctx.set_storage(storage)
    try:
        cases = [((41, 37, 201), dict(inline_padding=3, crossline_padding=4, angle_aperture=45.0)),
                 ((25, 23, 160), dict(inline_padding=2, crossline_padding=3, angle_aperture=30.0))]
        for shape, geo in cases:   # second call: a smaller survey
            print(f"noalloc {storage}: {shape} {geo}", flush=True)
            d_np, _ = synth.point_diffractor(*shape, noise=0.1, seed=11)
            data, vel = _device(d_np), _device(synth.production_velocity(*shape))
            prm = P.SurveyParams.for_cube(shape, **geo)
            info = P.survey_plan(prm, vel)
            torch.cuda.synchronize()
            ref = P.migrate_volume(data, vel, prm, ctx=ctx)
            out = torch.full((info.n_out_il, info.n_out_xl, info.n_t), float("nan"),
                             device="cuda:0")
            torch.cuda.synchronize()
            ctx.check(_noalloc(ctx, prm, data, vel, out), "psm_migrate_volume_noalloc")
            ctx.synchronize()
            assert torch.isfinite(out).all()
            assert torch.equal(out, ref), float((out - ref).abs().max())
            assert float(ref.abs().max()) > 0
    finally:
        ctx.set_storage("f32")


def test_noalloc_rejects_survey_beyond_workspace():
    import torch
    ctx = P.context(0)
    shape = (201, 201, 751)   # pair cube 245 MB > 64 MiB
    prm = P.SurveyParams.for_cube(shape, max_velocity=2200.0)
    small = torch.zeros(16, device="cuda:0")
    rc = _noalloc(ctx, prm, small, small, small)
    assert rc == PSM_ERR_ARG
    msg = ctx.lib.psm_last_error(ctx.h).decode()
    assert "workspace" in msg and str(64 << 20) in msg, msg
