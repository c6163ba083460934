"""Shared helpers for the parity tests: per-inline buffers built the way the
reference driver builds them (migration_kernel.cpp:385-511, main.cpp:329-352)."""
from __future__ import annotations

import numpy as np

import pyoracle as O


def window_call(data, vel, c, ap, pad_il, pad_xl):
    """Arguments of the reference kernel call for the output inline mapped to
    real inline c: (buf, vslice, il_pos, cur)."""
    n_il, n_xl, n_t = data.shape
    s_req = max(0, c - ap)
    e_req = min(n_il, c + ap + 1)
    s_buf = max(0, s_req - pad_il)
    e_buf = min(n_il, e_req + pad_il)
    K = n_xl + 2 * pad_xl
    buf = np.zeros((e_buf - s_buf, K, n_t), np.float32)
    buf[:, pad_xl:pad_xl + n_xl, :] = data[s_buf:e_buf]
    vs = np.zeros((K, n_t), np.float32)
    vs[pad_xl:pad_xl + n_xl] = vel[c]
    il_pos = np.arange(s_buf, e_buf, dtype=np.int32)
    cur = int(np.nonzero(il_pos == c)[0][0])
    return buf, vs, il_pos, cur


def t0_table(n_t, dt):
    return np.array([np.float32(i * dt) for i in range(n_t)], np.float32)


def reference_inline(buf, vs, il_pos, cur, xl_start, dt, t0, tan2, dx, dy, amp, n_xl=None):
    """The reference's own compiled kernel when oracle/_ref is built, else the
    oracle restatement."""
    impl = "ref" if O.have_ref() else "oracle"
    return O.kernel_inline(buf, vs, il_pos, cur, xl_start, dt, t0, tan2, dx, dy, amp,
                           n_xl=n_xl, impl=impl)


def reference_volume(data, vel, dt, dx, dy, pad_il, pad_xl, angle, amp, nthreads=0):
    if O.have_ref():
        return O.ref_migrate_volume(data, vel, dt, dx, dy, pad_il, pad_xl, angle, amp,
                                    nthreads=nthreads)
    return O.migrate_volume(data, vel, dt, dx, dy, pad_il, pad_xl, angle, amp, nthreads=nthreads)


rel_l2 = O.rel_l2

# per-trace bar: no single output trace may hide in a global relative L2
TRACE_TOL = 1e-4


def worst_trace_rel(got, ref):
    """Largest relative L2 error of one output trace (last axis = samples)
    among the traces where the reference is not zero; traces where the
    reference is exactly zero must be exactly zero."""
    g = np.asarray(got, np.float64).reshape(-1, np.shape(got)[-1])
    r = np.asarray(ref, np.float64).reshape(-1, np.shape(ref)[-1])
    den = np.linalg.norm(r, axis=1)
    zero = den == 0
    assert not g[zero].any(), "output nonzero where the reference is exactly zero"
    if zero.all():
        return 0.0
    return float((np.linalg.norm(g - r, axis=1)[~zero] / den[~zero]).max())
