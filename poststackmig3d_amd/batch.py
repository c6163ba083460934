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
from .kernel import SurveyParams, _ptr_where, PSM_DEVICE


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
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.psm_batch_create(self.ctx.h, arr, len(items), STORAGE[storage],
                                                     C.byref(h)), "psm_batch_create")
        self.h = h

    def set_shape(self, waves_per_block: int, blocks_per_cu: int = 0):
        self.ctx.check(self.ctx.lib.psm_batch_set_shape(self.h, waves_per_block, blocks_per_cu),
                       "psm_batch_set_shape")

    def autotune(self):
        w, b, ms = C.c_int(), C.c_int(), C.c_float()
        self.ctx.check(self.ctx.lib.psm_batch_autotune(self.h, C.byref(w), C.byref(b), C.byref(ms)),
                       "psm_batch_autotune")
        return w.value, b.value, ms.value

    def run(self, timed: bool = False):
        ms = C.c_float()
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
