"""Restatement of the reference's velocity providers (test infrastructure only).

Only tests/ import this module.  It turns a velocity source into the Vrms
cube [n_il][n_xl][n_t] on the data grid that the reference hands to its
kernel one inline at a time (MigrationKernel::prepareVelocitySliceWithPadding,
migration_kernel.cpp:470-511), and gives estimateMaxVelocity's probe values.

  segy_velocity_cube   SegyVelocityProvider::getVelocityIline
                       (velocity_reader.cpp:203-261): exact inline and
                       crossline with matching dt and n_t copy the trace
                       (:216-231); otherwise getVelocity / interpolateVelocity
                       (:165-201, :93-163): nearest labels, a 2x2 neighbourhood
                       (clamped), linear time interpolation per trace and an
                       inverse-(1 + |dil| + |dxl|) weighted mean in float.
  table_velocity_cube  TableVelocityProvider (:264-370): inverse-(1 + d)
                       weighting over all points, d the 4-D distance
                       (il, xl, t); exact match returns the point's value;
                       non-positive results fall back to the minimum.
  probe_max_velocity   main.cpp:22-64 through the provider's getVelocity.
"""
from __future__ import annotations

import bisect
import math

import numpy as np


def _f32(x):
    return float(np.float32(x))


class SegyVelocity:
    def __init__(self, cube, il_labels, xl_labels, dt):
        self.cube = np.asarray(cube, np.float32)      # [n_il][n_xl][n_t] on its own grid
        self.il = [int(v) for v in il_labels]
        self.xl = [int(v) for v in xl_labels]
        self.n_t = self.cube.shape[2]
        self.dt = float(dt)
        self.times = [i * self.dt for i in range(self.n_t)]
        self.il_idx = {v: i for i, v in enumerate(self.il)}
        self.xl_idx = {v: i for i, v in enumerate(self.xl)}

    def _trace(self, il, xl):
        # readTraceByCoords: missing trace -> zeros (segy_utils.cpp:223-241)
        if il in self.il_idx and xl in self.xl_idx:
            return self.cube[self.il_idx[il], self.xl_idx[xl]]
        return np.zeros(self.n_t, np.float32)

    def _time_interp(self, tr, time):
        # velocity_reader.cpp:124-138 / :183-196 (double arithmetic, float result)
        if time <= self.times[0]:
            return float(tr[0])
        if time >= self.times[-1]:
            return float(tr[-1])
        q = time / self.dt
        i = int(q)
        if i >= self.n_t - 1:
            return float(tr[-1])
        f = q - i
        return _f32(float(tr[i]) * (1.0 - f) + float(tr[i + 1]) * f)

    @staticmethod
    def _nearest(lst, v):
        # findNearestIndex (:70-91)
        k = bisect.bisect_left(lst, v)
        if k == 0:
            return 0
        if k == len(lst):
            return len(lst) - 1
        return k if abs(lst[k] - v) < abs(lst[k - 1] - v) else k - 1

    def interpolate(self, il, xl, time):
        # interpolateVelocity (:93-163)
        ii = self._nearest(self.il, il)
        xx = self._nearest(self.xl, xl)
        ils = [ii - 1 if ii > 0 else 0, ii + 1 if ii < len(self.il) - 1 else len(self.il) - 1]
        xls = [xx - 1 if xx > 0 else 0, xx + 1 if xx < len(self.xl) - 1 else len(self.xl) - 1]
        sv = np.float32(0.0)
        sw = np.float32(0.0)
        for a in ils:
            for b in xls:
                lv, xv = self.il[a], self.xl[b]
                vel = np.float32(self._time_interp(self._trace(lv, xv), time))
                w = np.float32(1.0) / (np.float32(1.0) + np.float32(abs(lv - il) + abs(xv - xl)))
                sv = np.float32(sv + vel * w)
                sw = np.float32(sw + w)
        return float(sv / sw) if sw > 0 else 0.0

    def get(self, il, xl, t_idx, dt):
        # getVelocity (:165-201)
        time = t_idx * dt
        if il in self.il_idx and xl in self.xl_idx:
            return self._time_interp(self._trace(il, xl), time)
        return self.interpolate(il, xl, time)

    def inline(self, il, xl_labels, n_t, dt):
        # getVelocityIline (:203-261)
        out = np.zeros((len(xl_labels), n_t), np.float32)
        exact_il = il in self.il_idx
        for k, xl in enumerate(xl_labels):
            if exact_il and xl in self.xl_idx:
                tr = self._trace(il, xl)
                if abs(dt - self.dt) < 1e-6 and n_t == self.n_t:
                    out[k, :min(n_t, tr.size)] = tr[:n_t]
                else:
                    for t in range(n_t):
                        out[k, t] = self.get(il, xl, t, dt)
            else:
                for t in range(n_t):
                    out[k, t] = self.interpolate(il, xl, t * dt)
        return out


class TableVelocity:
    def __init__(self, rows):
        self.pts = [(int(r[0]), int(r[1]), float(r[2]), _f32(r[3])) for r in rows]
        self.vmin = min(p[3] for p in self.pts)

    def get(self, il, xl, t_idx, dt):
        time = t_idx * dt
        sv = np.float32(0.0)
        sw = np.float32(0.0)
        for pil, pxl, pt, pv in self.pts:
            d = math.sqrt(float(abs(pil - il)) ** 2 + float(abs(pxl - xl)) ** 2 + abs(pt - time) ** 2)
            if d < 1e-10:
                return pv
            w = np.float32(1.0) / (np.float32(1.0) + np.float32(d))
            sv = np.float32(sv + np.float32(pv) * w)
            sw = np.float32(sw + w)
        r = float(sv / sw) if sw > 0 else self.vmin
        return self.vmin if r <= 0 else r

    def inline(self, il, xl_labels, n_t, dt):
        out = np.zeros((len(xl_labels), n_t), np.float32)
        for k, xl in enumerate(xl_labels):
            for t in range(n_t):
                out[k, t] = self.get(il, xl, t, dt)
        return out


class ConstantVelocity:
    def __init__(self, v):
        self.v = _f32(v)

    def get(self, il, xl, t_idx, dt):
        return self.v

    def inline(self, il, xl_labels, n_t, dt):
        return np.full((len(xl_labels), n_t), self.v, np.float32)


def velocity_cube(provider, il_labels, xl_labels, n_t, dt):
    """The Vrms cube on the data grid, one getVelocityIline per inline."""
    return np.stack([provider.inline(int(il), [int(x) for x in xl_labels], n_t, dt)
                     for il in il_labels])


def probe_max_velocity(provider, il_labels, xl_labels, n_t, dt):
    """estimateMaxVelocity (main.cpp:22-64)."""
    ils = [0] + ([len(il_labels) // 2, len(il_labels) - 1] if len(il_labels) > 1 else [])
    xls = [0] + ([len(xl_labels) // 2, len(xl_labels) - 1] if len(xl_labels) > 1 else [])
    vmax = np.float32(0.0)
    for a in ils:
        for b in xls:
            v = np.float32(provider.get(int(il_labels[a]), int(xl_labels[b]), n_t - 1, dt))
            if v > vmax:
                vmax = v
    return float(vmax) if vmax >= 100.0 else 3000.0
