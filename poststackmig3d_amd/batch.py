"""Batched migration of several resident volumes in one persistent,
graph-captured launch (psm_batch_* in include/psm.h).

    b = Batch(items, storage="f16")       # items: (SurveyParams, data, vel, out, alpha)
    b.autotune()                           # picks workgroup width x residency
    ms = b.run(timed=True)

Typical use is a velocity scan (the same cube and Vrms model under several
scales alpha) mixed with sub-volumes of other sizes; every output equals
migrate_volume of that item at velocity scale alpha.
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from ._lib import STORAGE, BatchItem, PsmError
from .kernel import SurveyParams, _ptr_where, _TorchOrder, PSM_DEVICE


class Batch:
    def __init__(self, items, storage: str = "f32", ctx: _lib.Context | None = None):
        self.ctx = ctx or _lib.context()
        if storage not in STORAGE:
            raise PsmError(f"unknown storage {storage!r}")
        arr = (BatchItem * len(items))()
        self._keep = []
        for k, (prm, data, vel, out, alpha) in enumerate(items):
            assert isinstance(prm, SurveyParams)
            ptrs = []
            for name, t in (("data", data), ("vel", vel), ("out", out)):
                p, w = _ptr_where(t, name)
                if w != PSM_DEVICE:
                    raise PsmError("batch items must be device tensors")
                ptrs.append(p)
            arr[k].survey = prm.to_c()
            self._keep.append(prm)
            arr[k].data, arr[k].vel, arr[k].out = ptrs
            arr[k].alpha = float(alpha)
        self._items = arr
        self._tensors = [t for it in items for t in it[1:4]]
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.psm_batch_create(self.ctx.h, arr, len(items), STORAGE[storage],
                                                     C.byref(h)), "psm_batch_create")
        self.h = h

    def set_shape(self, waves_per_block: int, blocks_per_cu: int = 0):
        self.ctx.check(self.ctx.lib.psm_batch_set_shape(self.h, waves_per_block, blocks_per_cu),
                       "psm_batch_set_shape")

    def shape(self):
        """(staged, waves_per_block, blocks_per_cu, steps) of the launch in use."""
        v = [C.c_int() for _ in range(4)]
        self.ctx.check(self.ctx.lib.psm_batch_get_shape(self.h, *[C.byref(x) for x in v]),
                       "psm_batch_get_shape")
        return bool(v[0].value), v[1].value, v[2].value, v[3].value

    def autotune(self):
        w, b, ms = C.c_int(), C.c_int(), C.c_float()
        self.ctx.check(self.ctx.lib.psm_batch_autotune(self.h, C.byref(w), C.byref(b), C.byref(ms)),
                       "psm_batch_autotune")
        return w.value, b.value, ms.value

    def run(self, timed: bool = False):
        """Replays the batch graph on the context's stream, ordered after the
        work torch has queued and before torch's later work."""
        ms = C.c_float()
        with _TorchOrder(self.ctx, *self._tensors[:1]):
            self.ctx.check(self.ctx.lib.psm_batch_run(self.h, C.byref(ms) if timed else None),
                           "psm_batch_run")
        return ms.value if timed else None

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.psm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


SCAN_ALPHAS = 16
SCAN_SUBVOLS = 8
SCAN_GEOMETRY = dict(dt=0.002, inline_step=25.0, crossline_step=25.0, inline_padding=0,
                     crossline_padding=0, angle_aperture=30.0, amp_correction=True)


def velocity_scan(device, n=(301, 301, 1001)):
    """The velocity-scan workload: the point-diffractor cube (V = 2000 m/s,
    25 Hz Ricker apex at the centre, 10 % noise; 30 deg, amplitude correction
    on, no padding) at n = 301 x 301 x 1001, migrated under 16 scales
    alpha = 0.85 .. 1.15 of its Vrms, plus 8 centred sub-volumes from
    64 x 64 x 501 to the whole cube at alpha = 1.  Returns (items, base,
    vbase): items = [(SurveyParams, data, vel, out, alpha)] with device
    tensors and empty outputs (the form Batch takes), base / vbase the host
    cube and velocity the items are cut from."""
    import numpy as np
    import torch

    from . import synth
    base_np, vbase_np = synth.point_diffractor(*n)
    base = torch.from_numpy(base_np).to(device)
    vbase = torch.from_numpy(vbase_np).to(device)

    def item(d, v, alpha):
        prm = SurveyParams.for_cube(tuple(d.shape), **SCAN_GEOMETRY)
        out = torch.empty(tuple(d.shape), dtype=torch.float32, device=device)  # no padding
        return (prm, d, v, out, float(alpha))

    items = [item(base, vbase, a) for a in np.linspace(0.85, 1.15, SCAN_ALPHAS)]
    sizes = np.linspace(64, n[0], SCAN_SUBVOLS).round().astype(int)
    lengths = np.linspace(501, n[2], SCAN_SUBVOLS).round().astype(int)
    for a, t in zip(sizes, lengths):
        o = (n[0] - a) // 2
        items.append(item(base[o:o + a, o:o + a, :t].contiguous(),
                          vbase[o:o + a, o:o + a, :t].contiguous(), 1.0))
    return items, base_np, vbase_np

