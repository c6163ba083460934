"""Batched, graph-captured persistent migration: a velocity scan over scaled
Vrms models plus uneven sub-volumes in one launch.  Each output must equal
the single-volume migration of that item at the same velocity scale by the
same form of the forward kernel in one launch (the same per-workgroup
arithmetic and summation order, so bit for bit), on every replay of the graph
with eager work in between, and after autotuning.  The batch takes its form
from the context's staging switch like psm_migrate_volume: staged by
default for every storage form, gather with staging off."""
import numpy as np
import pytest

import poststackmig3d_amd as P
from poststackmig3d_amd import synth
from poststackmig3d_amd._lib import Context

pytestmark = pytest.mark.gpu


def _items(torch, storage):
    base, vbase = synth.point_diffractor(41, 37, 180, v=2000.0, noise=0.1, seed=2)
    d = torch.from_numpy(base).cuda()
    v = torch.from_numpy(vbase).cuda()
    prm = P.SurveyParams.for_cube(d.shape, angle_aperture=30.0, inline_padding=2,
                                  crossline_padding=3)
    items = []
    for alpha in np.linspace(0.85, 1.15, 4):
        info = P.survey_plan(prm, v)
        out = torch.empty((info.n_out_il, info.n_out_xl, info.n_t), device="cuda")
        items.append((prm, d, v, out, float(alpha)))
    for (a, b, nt) in ((12, 9, 64), (23, 30, 150)):
        sd = d[:a, :b, :nt].contiguous()
        sv = torch.from_numpy(synth.production_velocity(a, b, nt)).cuda()
        sp = P.SurveyParams.for_cube(sd.shape, angle_aperture=45.0)
        info = P.survey_plan(sp, sv)
        out = torch.empty((info.n_out_il, info.n_out_xl, info.n_t), device="cuda")
        items.append((sp, sd, sv, out, 1.0))
    return items


def _reference(items, storage, wpb, staged):
    ctx = Context(0)
    try:
        ctx.set_storage(storage)
        if staged:
            # the batch's staged form: 8 crosslines, 4 offsets per group, one launch
            ctx.set_tuning(8, 0)
            ctx.set_staging(True, 4)
        else:
            ctx.set_tuning(wpb, 0)
            ctx.set_staging(False)
        refs = []
        for prm, d, v, out, alpha in items:
            ctx.set_velocity_scale(alpha, alpha)
            refs.append(P.migrate_volume(d, v, prm, ctx=ctx).cpu().numpy())
        return refs
    finally:
        ctx.close()


@pytest.mark.parametrize("storage,staging",
                         [("f32", None), ("f16", None), ("bf16", None), ("f32", False),
                          ("f16", False)])
def test_batch_matches_single_volumes(storage, staging):
    import torch
    from poststackmig3d_amd.batch import Batch
    items = _items(torch, storage)
    staged = staging is not False
    refs = _reference(items, storage, 4, staged)
    bctx = Context(0)
    bctx.set_staging(staging)
    b = Batch(items, storage=storage, ctx=bctx)
    try:
        for replay in range(3):
            for _, _, _, out, _ in items:
                out.fill_(float("nan"))
            torch.cuda.synchronize()
            b.run()
            b.ctx.synchronize()
            for k, (_, _, _, out, _) in enumerate(items):
                assert np.array_equal(out.cpu().numpy(), refs[k]), (replay, k)
            # eager work between replays
            P.migrate_volume(items[0][1], items[0][2], items[0][0])
        w, bpc, ms = b.autotune()
        assert w in (1, 2, 4, 8) and ms > 0
        is_staged, w2, bpc2, steps = b.shape()
        assert is_staged == staged and w2 == w and bpc2 == bpc
        assert steps in ((2, 4) if staged else (0,))
        assert not staged or w in (4, 8)
        # every staged shape gives the same result: the forced shapes of the tuner
        if staged:
            for ww, gg in ((4, 2), (8, 2), (4, 4)):
                b.ctx.set_staging(True, gg)
                b.set_shape(ww, 0)
                assert b.shape()[1:] == (ww, 5 if ww == 4 else 3, gg)
                for _, _, _, out, _ in items:
                    out.fill_(float("nan"))
                torch.cuda.synchronize()
                b.run()
                b.ctx.synchronize()
                for k, (_, _, _, out, _) in enumerate(items):
                    assert np.array_equal(out.cpu().numpy(), refs[k]), (ww, gg, k)
            b.ctx.set_staging(staging)
            b.set_shape(w, bpc)
        b.run()
        b.ctx.synchronize()
        for k, (_, _, _, out, _) in enumerate(items):
            assert np.array_equal(out.cpu().numpy(), refs[k])
    finally:
        b.close()
        bctx.close()
