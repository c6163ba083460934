"""Pins the oracle (CPU restatement of the reference path) before it is trusted.

  * golden vectors written by the reference CLI (tests/golden/, six cases:
    constant / SEG-Y / table velocity, padding, IBM float, shuffled and
    missing traces, resampled velocity grid)
  * the reference's own kernel compiled from its sources (oracle/_ref), where
    built
  * closed forms and invariants: the zero-offset term, a single spike's
    hyperbola, constant-data pair counts, linearity, mirror symmetry, t0 = 0
    and leading-padding zeros, FP32 vs FP64 agreement
"""
import numpy as np
import pytest

import pyoracle as O
from _golden import GOLDEN, inputs, load, output

from poststackmig3d_amd import segy, synth

DT = 0.002


def _migrate(cube, vel, dt, num, ill, precision="f32"):
    return O.migrate_volume(cube, vel, dt, num["inline_step"], num["crossline_step"],
                            num["inline_padding"], num["crossline_padding"],
                            num["angle_aperture"], num["amp_correction"], precision=precision,
                            il_labels=ill)


@pytest.mark.parametrize("path", GOLDEN, ids=lambda p: p.split("/")[-1][:-4])
def test_oracle_reproduces_reference_cli(path):
    """Oracle restatement vs the reference binary's output file."""
    fx = load(path)
    cube, vel, ill, xll, dt, num, vmax, seg = inputs(fx)
    ref, oil, oxl = output(fx)
    got = _migrate(cube, vel, dt, num, ill)
    assert got.shape == ref.shape
    if seg["fmt"] == 1:
        # the reference writes IBM floats, truncating the fraction
        got = segy.ibm_to_ieee(segy.ieee_to_ibm(got)).reshape(got.shape)
    err = O.rel_l2(got, ref)
    assert err <= 2e-6, (fx["name"], err)
    # geometry printed by the reference (main.cpp:204-210)
    t2, ap = O.aperture(num["angle_aperture"], vmax, cube.shape[2], dt, num["inline_step"])
    assert f"IL=±{ap} lines" in str(fx["stdout"])
    assert np.array_equal(oil, O.output_labels(ill, num["inline_padding"]))
    assert np.array_equal(oxl, O.output_labels(xll, num["crossline_padding"]))


@pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref not built (no reference sources)")
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_oracle_matches_reference_library(seed):
    """Random single calls: odd il_pos, xl offsets, zero-velocity holes."""
    rng = np.random.default_rng(seed)
    nb, K, n_t = rng.integers(1, 9), rng.integers(3, 17), rng.integers(20, 90)
    buf = rng.normal(size=(nb, K, n_t)).astype(np.float32)
    vs = rng.uniform(1500, 3000, size=(K, n_t)).astype(np.float32)
    vs[rng.random((K, n_t)) < 0.2] = 0.0
    il_pos = np.sort(rng.choice(40, nb, replace=False)).astype(np.int32)
    cur = int(rng.integers(-1, nb + 1))
    xl0 = int(rng.integers(-2, 3))
    n_xl = int(rng.integers(1, K + 3))
    t0 = np.array([np.float32(i * DT) for i in range(n_t)], np.float32)
    tan2 = float(np.float32(np.tan(np.deg2rad(rng.uniform(20, 70)))) ** 2)
    dx, dy = float(rng.uniform(10, 40)), float(rng.uniform(10, 40))
    for amp in (True, False):
        a = O.kernel_inline(buf, vs, il_pos, cur, xl0, DT, t0, tan2, dx, dy, amp, n_xl=n_xl)
        b = O.kernel_inline(buf, vs, il_pos, cur, xl0, DT, t0, tan2, dx, dy, amp, n_xl=n_xl,
                            impl="ref")
        assert O.rel_l2(a, b) <= 1e-6


def test_zero_offset_term():
    """One aperture trace at the image point: tau = t0, so out[s] is the
    sample itself times 1/(t0 v^2) (migration_kernel.cpp:266-271)."""
    rng = np.random.default_rng(4)
    n_t, v = 64, 2300.0
    x = rng.normal(size=(1, 1, n_t)).astype(np.float32)
    vs = np.full((1, n_t), v, np.float32)
    t0 = np.array([np.float32(i * DT) for i in range(n_t)], np.float32)
    for amp in (False, True):
        out = O.kernel_inline(x, vs, [0], 0, 0, DT, t0, 0.33, 25.0, 25.0, amp)[0]
        with np.errstate(divide="ignore"):
            w = np.ones(n_t) if not amp else 1.0 / (np.arange(n_t) * DT * v * v)
        s = np.arange(1, n_t - 1)
        assert out[0] == 0.0
        np.testing.assert_allclose(out[s] / w[s], x[0, 0, s], rtol=2e-4, atol=2e-4)


def test_single_spike_hyperbola():
    """A unit spike at sample s0 of trace (b0, k0), amplitude correction off:
    an image sample responds with (1 - f) when its travel time to that trace
    falls on sample s0 (i = s0) and f when it falls on s0 - 1, and only inside
    the aperture r <= (v t0 / 2) tan(theta); evaluated independently in
    float64 here, with dx != dy so the axes cannot be swapped."""
    n_il, n_xl, n_t = 9, 11, 120
    dx, dy, v, ang = 20.0, 30.0, 1800.0, 50.0
    b0, k0, s0 = 3, 7, 70
    data = np.zeros((n_il, n_xl, n_t), np.float32)
    data[b0, k0, s0] = 1.0
    vel = np.full(data.shape, v, np.float32)
    out = O.migrate_volume(data, vel, DT, dx, dy, 0, 0, ang, False, precision="f32")
    T2 = np.tan(np.deg2rad(ang)) ** 2
    exp = np.zeros_like(out, dtype=np.float64)
    for c in range(n_il):
        for j in range(n_xl):
            r2 = ((b0 - c) * dx) ** 2 + ((k0 - j) * dy) ** 2
            for s in range(1, n_t):
                t0 = s * DT
                if r2 > (v * t0 / 2) ** 2 * T2:
                    continue
                q = np.sqrt(t0 * t0 + 4 * r2 / v / v) / DT
                i = int(q)
                f = q - i
                if i == s0:
                    exp[c, j, s] = 1 - f
                elif i == s0 - 1:
                    exp[c, j, s] = f
    # aperture edges and integer crossings are the only places float32 and
    # float64 may disagree; they are few
    bad = np.abs(out - exp) > 2e-3
    assert bad.sum() <= 3, np.argwhere(bad)[:5]
    assert np.abs(out).max() <= 1.0 + 1e-6 and out.min() >= -1e-6
    assert (out > 0).sum() > 50


def test_constant_data_counts_gathering_pairs():
    """Constant unit data, amplitude correction off: every gathering pair adds
    exactly 1 (lerp of equal samples), so the image sums to P_g; checks the
    kernel restatement against the independent pair counter."""
    n = (13, 12, 90)
    vel = synth.production_velocity(*n)
    data = np.ones(n, np.float32)
    out = O.migrate_volume(data, vel, DT, 25.0, 25.0, 2, 3, 45.0, False)
    assert np.array_equal(out, np.round(out))
    p_ap, p_g = O.count_pairs(vel, DT, 25.0, 25.0, 2, 3, 45.0)
    assert float(out.astype(np.float64).sum()) == p_g
    assert p_ap >= p_g > 0


def test_linearity_and_mirror_symmetry():
    rng = np.random.default_rng(7)
    n = (11, 11, 70)
    x = rng.normal(size=n).astype(np.float32)
    y = rng.normal(size=n).astype(np.float32)
    vel = np.full(n, 2100.0, np.float32)
    m = lambda d: O.migrate_volume(d, vel, DT, 25.0, 25.0, 0, 0, 40.0, True, precision="f64")
    lhs = m((2 * x + 3 * y).astype(np.float32))
    rhs = 2 * m(x) + 3 * m(y)
    assert O.rel_l2(lhs, rhs) <= 1e-6
    # mirror along the inline axis: constant velocity, mirrored data
    xm = x[::-1].copy()
    assert O.rel_l2(m(xm), m(x)[::-1]) <= 1e-12


def test_zero_rows_and_fp64_agreement():
    n = (15, 13, 80)
    data, _ = synth.point_diffractor(*n, noise=0.2, seed=3, t_apex=0.08)
    vel = synth.production_velocity(*n)
    a = O.migrate_volume(data, vel, DT, 25.0, 25.0, 2, 3, 45.0, True)
    b = O.migrate_volume(data, vel, DT, 25.0, 25.0, 2, 3, 45.0, True, precision="f64")
    assert not a[:, :, 0].any()          # t0 = 0 row (migration_kernel.cpp:121)
    assert not a[:, :3].any()            # leading padded crosslines: no valid sample yet
    assert np.abs(a[:, -3:]).max() > 0   # trailing padded crosslines: sticky parameters
    assert np.array_equal(a[0], a[2]) and np.array_equal(a[-1], a[-3])  # padded inlines
    assert O.rel_l2(a, b) <= 1e-4


def test_geometry_helpers():
    # main.cpp:144-172 output labels and :258-288 nearest-inline mapping
    lab = np.array([10, 12, 14, 20], np.int32)
    assert list(O.output_labels(lab, 2)) == [6, 8, 10, 12, 14, 16, 18, 20, 22, 24]
    assert [O.map_output_inline(lab, l) for l in (6, 16, 17, 18, 24)] == [0, 2, 2, 3, 3]
    # main.cpp:22-64: probe at 9 points of the last sample, 3000 below 100 m/s
    v = np.full((5, 4, 10), 50.0, np.float32)
    assert O.estimate_max_velocity(v) == 3000.0
    v[4, 2, 9] = 2500.0
    assert O.estimate_max_velocity(v) == 2500.0
    v[1, 1, 9] = 9000.0                  # not a probe point
    assert O.estimate_max_velocity(v) == 2500.0
    # plain-method case: V = 2000, 1 s, 30 deg, 25 m -> window +-24 inlines
    t2, ap = O.aperture(30.0, 2000.0, 501, DT, 25.0)
    assert ap == 24 and abs(t2 - 1.0 / 3.0) < 1e-7
