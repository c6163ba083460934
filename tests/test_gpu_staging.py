"""The LDS-staged forward kernel against the gather kernel and the reference.

Both kernel forms evaluate the same terms (same mask, travel time, sample
index, interpolation and weight per pair) and differ only in FP32 summation
order, so they agree to rounding (rel-L2 <= 2e-6 here); each matches the
reference kernel to the parity bar (1e-5).  The cases reach every path of the
staged kernel: sticky trailing crosslines (the workgroup runs the gather
path), coarse crossline bins with a slow velocity (time windows wider than a
staged slot: per-row fallback), fine crossline bins (more column offsets than
the column table holds: per-row fallback), traces shorter than a slot, large
samples at both ends of every trace (windows that run past the trace end),
and both storage forms.
"""
import numpy as np
import pytest

from _cases import TRACE_TOL, reference_volume, rel_l2, worst_trace_rel

import poststackmig3d_amd as P
from poststackmig3d_amd import synth
from poststackmig3d_amd._lib import Context, PsmError

pytestmark = pytest.mark.gpu

TOL_REF = 1e-5
TOL_FORMS = 2e-6


def _case(name):
    if name == "padded_lateral":
        n = (29, 31, 140)
        data, _ = synth.production(*n, n_diffractors=4, seed=8)
        vel = synth.production_velocity(*n)
        geo = dict(inline_step=25.0, crossline_step=25.0, inline_padding=3, crossline_padding=4,
                   angle_aperture=45.0)
    elif name == "point_unpadded":
        n = (41, 45, 160)
        data, vel = synth.point_diffractor(*n, noise=0.1, seed=4)
        geo = dict(inline_step=25.0, crossline_step=25.0, angle_aperture=30.0)
    elif name == "coarse_slow":
        # 200 m crossline bins at 1000 m/s: one crossline step moves the travel
        # time by up to 2 sin(70) 200 / 1000 s = 188 samples, so the windows of
        # columns read by several waves exceed a slot
        n = (15, 17, 600)
        data, _ = synth.point_diffractor(*n, dy=200.0, v=1000.0, noise=0.2, seed=5)
        vel = np.full(n, 1000.0, np.float32)
        geo = dict(inline_step=25.0, crossline_step=200.0, crossline_padding=2,
                   angle_aperture=70.0)
    elif name == "fine_bins":
        # 0.5 m crossline bins: the aperture spans more crossline offsets than
        # the column table holds
        n = (3, 1100, 300)
        data, _ = synth.point_diffractor(*n, dy=0.5, noise=0.1, seed=6)
        vel = np.full(n, 2000.0, np.float32)
        geo = dict(inline_step=25.0, crossline_step=0.5, angle_aperture=60.0)
    elif name == "mid_bins":
        # 100 m crossline bins at 2000 m/s, 60 deg: the windows of the first
        # columns of a row fit a slot while columns further out exceed it, so
        # the row's fallback vote must see the columns of every lane
        n = (9, 60, 600)
        data, _ = synth.point_diffractor(*n, dy=100.0, v=2000.0, noise=0.2, seed=7)
        vel = np.full(n, 2000.0, np.float32)
        geo = dict(inline_step=25.0, crossline_step=100.0, angle_aperture=60.0)
    elif name == "short_traces":
        # 40 samples: every window reaches the trace end, and the staging
        # loads stop at it (the lane holding the last pair reads into the next
        # trace of the row)
        n = (21, 23, 40)
        data, _ = synth.production(*n, n_diffractors=3, seed=9)
        vel = synth.production_velocity(*n)
        geo = dict(inline_step=25.0, crossline_step=25.0, inline_padding=2, crossline_padding=3,
                   angle_aperture=60.0)
    elif name in ("end_spikes_even", "end_spikes_odd"):
        # every trace starts with a large sample and ends with two: a term
        # that read past the last pair (the next trace's first pair, or a
        # pair the bound i < n_t - 1 excludes) would stand out far above the
        # bar; n_t - 1 even and odd, since the staged windows that run past
        # the trace end are aligned on it
        n = (19, 21, 301 if name == "end_spikes_even" else 300)
        data, _ = synth.production(*n, n_diffractors=3, seed=10)
        data[:, :, 0] = 1000.0
        data[:, :, -1] = 100.0
        data[:, :, -2] = -50.0
        vel = synth.production_velocity(*n)
        geo = dict(inline_step=25.0, crossline_step=25.0, inline_padding=2, crossline_padding=3,
                   angle_aperture=45.0)
    else:
        raise KeyError(name)
    prm = P.SurveyParams.for_cube(n, **geo)
    ref_args = (0.002, geo["inline_step"], geo["crossline_step"], geo.get("inline_padding", 0),
                geo.get("crossline_padding", 0), geo["angle_aperture"], True)
    return data, vel, prm, ref_args


CASE_NAMES = ["padded_lateral", "point_unpadded", "coarse_slow", "mid_bins", "fine_bins",
              "short_traces", "end_spikes_even", "end_spikes_odd"]


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def _migrate(ctx, data, vel, prm, staging, storage="f32", wpb=4, steps=4, rows=-1):
    ctx.set_storage(storage)
    ctx.set_tuning(wpb, rows)
    ctx.set_staging(staging, steps)
    try:
        return P.migrate_volume(data, vel, prm, ctx=ctx)
    finally:
        ctx.set_storage("f32")
        ctx.set_tuning(4, -1)
        ctx.set_staging(None, 4)


@pytest.mark.parametrize("name", CASE_NAMES)
def test_staged_matches_reference_and_gather(ctx, name):
    data, vel, prm, ref_args = _case(name)
    print(f"staging case {name} {data.shape}", flush=True)
    ref = reference_volume(data, vel, *ref_args)
    gather = _migrate(ctx, data, vel, prm, staging=False)
    assert rel_l2(gather, ref) <= TOL_REF
    for wpb in (4, 8):
        for steps in (2, 4):
            got = _migrate(ctx, data, vel, prm, staging=True, wpb=wpb, steps=steps)
            assert np.isfinite(got).all()
            e_ref, e_g = rel_l2(got, ref), rel_l2(got, gather)
            worst = worst_trace_rel(got, ref)
            print(f"  wpb {wpb} steps {steps}: vs reference {e_ref:.2e} (worst trace "
                  f"{worst:.2e}), vs gather {e_g:.2e}", flush=True)
            assert e_ref <= TOL_REF, (wpb, steps, e_ref)
            assert worst <= TRACE_TOL, (wpb, steps, worst)
            assert e_g <= TOL_FORMS, (wpb, steps, e_g)


@pytest.mark.parametrize("name", CASE_NAMES)
def test_staged_adjoint_matches_gather_adjoint(ctx, name):
    """The staged adjoint (the workgroup's image samples scattered into LDS
    windows, each window flushed once into a row-pair buffer that a second
    kernel adds into both data rows of the pair) against the gather-form
    adjoint (every pair scattered with global atomics): the same terms summed
    in another order.  The rows whose windows or offsets do not fit
    (coarse_slow, fine_bins) and sticky workgroups take the gather-form
    adjoint inside the staged kernel.  Passes of 1 and 3 row offsets (the
    row-pair buffer holding part of the aperture at a time) give the same
    data to rounding."""
    data, vel, prm, _ = _case(name)
    info = P.survey_plan(prm, vel)
    rng = np.random.default_rng(11)
    image = rng.standard_normal((info.n_out_il, info.n_out_xl, info.n_t)).astype(np.float32)
    print(f"adjoint case {name} image {image.shape}", flush=True)
    ctx.set_staging(False)
    try:
        gather = P.demigrate_volume(image, vel, prm, ctx=ctx)
    finally:
        ctx.set_staging(None, 4)
    staged = P.demigrate_volume(image, vel, prm, ctx=ctx)
    assert np.isfinite(staged).all()
    e = rel_l2(staged, gather)
    print(f"  staged vs gather adjoint {e:.2e}", flush=True)
    assert e <= TOL_FORMS, e
    for rows in (1, 3):
        ctx.set_tuning(4, rows)
        try:
            passes = P.demigrate_volume(image, vel, prm, ctx=ctx)
        finally:
            ctx.set_tuning(4, -1)
        e = rel_l2(passes, gather)
        print(f"  staged adjoint in passes of {rows} row offsets vs gather {e:.2e}", flush=True)
        assert e <= TOL_FORMS, (rows, e)


@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("name", CASE_NAMES)
def test_staged_half_storage_matches_gather(ctx, name, storage):
    """Half-precision traces: the staged form widens each stored pair, scales
    it by its trace's scale and sums the rows +db and -db into one FP32
    window; the gather form interpolates each stored pair and scales the
    result.  The same terms in another FP32 order: they agree to rounding."""
    data, vel, prm, _ = _case(name)
    print(f"staging case {name} {data.shape} {storage}", flush=True)
    gather = _migrate(ctx, data, vel, prm, staging=False, storage=storage)
    for wpb, steps in ((4, 4), (8, 2), (8, 4)):
        got = _migrate(ctx, data, vel, prm, staging=True, storage=storage, wpb=wpb, steps=steps)
        assert np.isfinite(got).all()
        e, worst = rel_l2(got, gather), worst_trace_rel(got, gather)
        print(f"  {name} {storage} wpb {wpb} steps {steps}: vs gather {e:.2e} (worst trace "
              f"{worst:.2e})", flush=True)
        assert e <= TOL_FORMS, (wpb, steps, e)
        assert worst <= TRACE_TOL, (wpb, steps, worst)


def test_staged_row_passes(ctx):
    """Aperture-row passes (later launches accumulate) with the staged form."""
    data, vel, prm, ref_args = _case("padded_lateral")
    ref = reference_volume(data, vel, *ref_args)
    for rows in (0, 1, 3):
        got = _migrate(ctx, data, vel, prm, staging=True, rows=rows)
        assert rel_l2(got, ref) <= TOL_REF, rows


def test_staging_switch_roundtrip(ctx):
    assert ctx.get_staging() == (None, 4)  # default: chosen by trace storage
    ctx.set_staging(False, 2)
    assert ctx.get_staging() == (False, 2)
    ctx.set_staging(True, 4)
    assert ctx.get_staging() == (True, 4)
    with pytest.raises(PsmError):
        ctx.set_staging(True, 3)
    ctx.set_staging(None, 4)


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_automatic_form_matches_forced_forms(ctx, storage):
    """The default (automatic) form equals the form it selects for every
    storage form: staged with 8 crosslines per workgroup and 4 offsets per
    group, the whole aperture in one launch."""
    data, vel, prm, _ = _case("padded_lateral")
    auto = _migrate(ctx, data, vel, prm, staging=None, storage=storage)
    forced = _migrate(ctx, data, vel, prm, staging=True, storage=storage, wpb=8, steps=4, rows=0)
    assert np.array_equal(auto, forced)
