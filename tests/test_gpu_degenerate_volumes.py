"""Degenerate survey shapes through the whole-volume call in its default
(staged) form: a single trace, a single inline or crossline (with and
without padding), one to three samples, and tiles of exactly 64 and 65
samples.  One inline or crossline label gives main.cpp no label step, so no
padded output lines on that axis (main.cpp:144-171), while the kernel's
buffer is padded all the same (main.cpp:296-307): with one crossline and
crossline padding, the one output crossline falls on a zero-velocity padding
trace and the image is zero.  FP32 traces against the compiled reference kernel driven like
main.cpp (rel-L2 <= 1e-5, per output trace <= 1e-4, exact zeros where the
reference is zero); FP16 / BF16 staged against the gather form of the same
storage (rel-L2 <= 2e-6)."""
import numpy as np
import pytest

from _cases import TRACE_TOL, reference_volume, rel_l2, worst_trace_rel

import poststackmig3d_amd as P
from poststackmig3d_amd._lib import Context

pytestmark = pytest.mark.gpu

SHAPES = [((1, 1, 50), (0, 0)), ((1, 1, 50), (2, 3)), ((1, 9, 30), (0, 2)),
          ((9, 1, 30), (3, 0)), ((3, 4, 1), (1, 1)), ((3, 4, 2), (0, 0)),
          ((5, 6, 3), (2, 2)), ((2, 2, 64), (0, 0)), ((4, 5, 65), (1, 1)),
          ((2, 1, 40), (1, 2))]


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("shape,pad", SHAPES, ids=[f"{s}-pad{p}" for s, p in SHAPES])
def test_degenerate_volume(ctx, shape, pad):
    rng = np.random.default_rng(sum(shape) * 31 + pad[0] * 7 + pad[1])
    data = rng.standard_normal(shape).astype(np.float32)
    vel = (2000.0 + 300.0 * rng.random(shape)).astype(np.float32)
    geo = dict(dt=0.002, inline_step=25.0, crossline_step=25.0, inline_padding=pad[0],
               crossline_padding=pad[1], angle_aperture=45.0, amp_correction=True)
    print(f"degenerate volume {shape} pad {pad}", flush=True)
    ref = reference_volume(data, vel, geo["dt"], geo["inline_step"], geo["crossline_step"],
                           pad[0], pad[1], geo["angle_aperture"], True)
    prm = P.SurveyParams.for_cube(shape, **geo)
    got = P.migrate_volume(data, vel, prm, ctx=ctx)
    assert got.shape == ref.shape
    assert np.isfinite(got).all()
    if ref.any():
        assert rel_l2(got, ref) <= 1e-5, rel_l2(got, ref)
        assert worst_trace_rel(got, ref) <= TRACE_TOL
    else:
        assert not got.any()
    for storage in ("f16", "bf16"):
        ctx.set_storage(storage)
        try:
            staged = P.migrate_volume(data, vel, prm, ctx=ctx)
            ctx.set_staging(False)
            gather = P.migrate_volume(data, vel, prm, ctx=ctx)
        finally:
            ctx.set_staging(None, 4)
            ctx.set_storage("f32")
        assert np.isfinite(staged).all()
        if gather.any():
            assert rel_l2(staged, gather) <= 2e-6, (storage, rel_l2(staged, gather))
        else:
            assert not staged.any()
