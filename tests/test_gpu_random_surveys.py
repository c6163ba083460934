"""Seeded random surveys through the default (staged) forward form: random cube
shapes, bin sizes, paddings, apertures, amplitude correction, and a laterally
and temporally varying Vrms with zero-velocity holes (the reference's sticky
per-sample parameters).  Each survey is migrated by the default form of every
trace storage:

  FP32 traces   against the compiled reference kernel driven like main.cpp
                (rel-L2 <= 1e-5, worst output trace <= 1e-4, exact zeros where
                the reference is zero)
  FP16 / BF16   against the gather form of the same storage (the same terms in
                another FP32 order: rel-L2 <= 2e-6, worst trace <= 1e-4)

The shapes reach short rows (a few column offsets), rows longer than the
first fill, traces shorter than a staged window, and windows that run past
the trace end."""
import numpy as np
import pytest

from _cases import TRACE_TOL, reference_volume, rel_l2, worst_trace_rel

import poststackmig3d_amd as P
from poststackmig3d_amd._lib import Context

pytestmark = pytest.mark.gpu

N_SURVEYS = 24
TOL_REF = 1e-5
TOL_FORMS = 2e-6


def _survey(seed):
    rng = np.random.default_rng(1000 + seed)
    n = (int(rng.integers(7, 34)), int(rng.integers(7, 41)), int(rng.integers(40, 301)))
    dt = float(rng.choice([0.001, 0.002, 0.004]))
    dx = float(rng.choice([12.5, 25.0, 37.5, 50.0]))
    dy = float(rng.choice([12.5, 25.0, 37.5, 50.0]))
    geo = dict(dt=dt, inline_step=dx, crossline_step=dy,
               inline_padding=int(rng.integers(0, 6)), crossline_padding=int(rng.integers(0, 6)),
               angle_aperture=float(rng.uniform(20.0, 70.0)),
               amp_correction=bool(rng.integers(0, 2)))
    data = rng.standard_normal(n).astype(np.float32)
    # a few strong samples, at both trace ends too
    data[:, :, 0] += 20.0 * rng.standard_normal(n[:2]).astype(np.float32)
    data[:, :, -1] += 20.0 * rng.standard_normal(n[:2]).astype(np.float32)
    il = np.arange(n[0])[:, None, None]
    xl = np.arange(n[1])[None, :, None]
    t = (np.arange(n[2]) * dt)[None, None, :]
    v0 = float(rng.uniform(1200.0, 3500.0))
    vel = (v0 + float(rng.uniform(0.0, 1500.0)) * t
           + 200.0 * np.sin(2 * np.pi * il / max(n[0], 2)) * np.cos(2 * np.pi * xl / max(n[1], 2)))
    vel = vel.astype(np.float32)
    # zero-velocity holes: whole traces and sample ranges
    holes = rng.random(n[:2]) < 0.05
    vel[holes] = 0.0
    k = int(rng.integers(0, n[1]))
    vel[:, k, : n[2] // 3] = 0.0
    return data, vel, geo


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("seed", range(N_SURVEYS))
def test_random_survey_default_forms(ctx, seed):
    data, vel, geo = _survey(seed)
    prm = P.SurveyParams.for_cube(data.shape, **geo)
    print(f"survey {seed}: {data.shape} dt {geo['dt']} bins ({geo['inline_step']}, "
          f"{geo['crossline_step']}) pad ({geo['inline_padding']}, {geo['crossline_padding']}) "
          f"{geo['angle_aperture']:.1f} deg amp {geo['amp_correction']}", flush=True)
    ref = reference_volume(data, vel, geo["dt"], geo["inline_step"], geo["crossline_step"],
                           geo["inline_padding"], geo["crossline_padding"],
                           geo["angle_aperture"], geo["amp_correction"])
    got = P.migrate_volume(data, vel, prm, ctx=ctx)
    assert np.isfinite(got).all()
    e, worst = rel_l2(got, ref), worst_trace_rel(got, ref)
    print(f"  f32 default vs reference {e:.2e} (worst trace {worst:.2e})", flush=True)
    assert e <= TOL_REF, e
    assert worst <= TRACE_TOL, worst
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
        e, worst = rel_l2(staged, gather), worst_trace_rel(staged, gather)
        print(f"  {storage} default vs gather {e:.2e} (worst trace {worst:.2e})", flush=True)
        assert e <= TOL_FORMS, (storage, e)
        assert worst <= TRACE_TOL, (storage, worst)
