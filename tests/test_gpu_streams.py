"""Ordering against torch streams: inputs written by torch work that is still
queued on another stream when the call is made.  The Python wrappers make
the context's stream wait for torch's current stream before the call and
torch's stream wait for the context's after it, and the volume calls probe
the maximum velocity on the context's stream, so the results equal those on
inputs that were complete -- with the maximum velocity passed explicitly
(no probe) and probed."""
import numpy as np
import pytest

import poststackmig3d_amd as P
from poststackmig3d_amd import synth
from poststackmig3d_amd._lib import Context

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("explicit_vmax", [True, False])
def test_inputs_from_a_busy_side_stream(explicit_vmax):
    import torch
    n = (41, 37, 201)
    d_np, _ = synth.point_diffractor(*n, noise=0.1, seed=9)
    v_np = synth.production_velocity(*n)
    prm = P.SurveyParams.for_cube(n, inline_padding=3, crossline_padding=4, angle_aperture=45.0)
    if explicit_vmax:
        prm.max_velocity = float(v_np[..., -1].max())
    ref = P.migrate_volume(d_np, v_np, prm)            # host path: synchronous
    d0 = torch.from_numpy(d_np).cuda()
    v0 = torch.from_numpy(v_np).cuda()
    torch.cuda.synchronize()
    ctx = Context(0)                                    # a stream of its own
    side = torch.cuda.Stream()
    try:
        with torch.cuda.stream(side):
            # tens of milliseconds of work queued ahead of the inputs
            x = torch.randn(4096, 4096, device="cuda")
            for _ in range(12):
                x = torch.tanh(torch.mm(x, x) * (1.0 / 64.0))
            data = d0 * 1.0 + x[0, 0] * 0.0
            vel = v0 * 1.0
            img = P.migrate_volume(data, vel, prm, ctx=ctx, sync=False)
            got = img.cpu().numpy()                     # ordered after the call
        assert np.array_equal(got, ref)
        # and the batch / dot paths on the same busy stream
        with torch.cuda.stream(side):
            for _ in range(12):
                x = torch.tanh(torch.mm(x, x) * (1.0 / 64.0))
            y = img * 2.0 + x[0, 0] * 0.0
            dot = P.dot_f64(y, img, ctx=ctx)
        ref_dot = 2.0 * float(np.dot(ref.ravel().astype(np.float64), ref.ravel()))
        assert abs(dot - ref_dot) <= 1e-9 * abs(ref_dot)
    finally:
        torch.cuda.synchronize()
        ctx.close()
