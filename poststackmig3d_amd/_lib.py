"""ctypes binding of libpsm_hip.so (the C ABI declared in include/psm.h).

The library is built in-tree by __graft_entry__.build() (make -C
poststackmig3d_amd/csrc).  There is no CPU fallback: if the library is missing
or no gfx950 device is usable, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsm_hip.so")

PSM_OK = 0
PSM_HOST = 0
PSM_DEVICE = 1
STORAGE = {"f32": 0, "f16": 1, "bf16": 2}

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)


class Survey(C.Structure):
    """psm_survey (include/psm.h)."""
    _fields_ = [
        ("n_il", C.c_size_t), ("n_xl", C.c_size_t), ("n_t", C.c_size_t),
        ("dt", C.c_double), ("inline_step", C.c_double), ("crossline_step", C.c_double),
        ("inline_padding", C.c_int), ("crossline_padding", C.c_int),
        ("angle_aperture", C.c_double), ("amp_correction", C.c_int),
        ("inline_labels", _ip), ("crossline_labels", _ip),
        ("max_velocity", C.c_float),
    ]


class Plan(C.Structure):
    """psm_plan (include/psm.h)."""
    _fields_ = [
        ("n_out_il", C.c_size_t), ("n_out_xl", C.c_size_t), ("n_t", C.c_size_t),
        ("max_aperture_il", C.c_int), ("tan_theta_sq", C.c_float), ("max_velocity", C.c_float),
        ("out_inline_labels", _ip), ("out_crossline_labels", _ip), ("out_to_in_il", _ip),
        ("window", _ip), ("t0_times", _fp),
    ]


class BatchItem(C.Structure):
    """psm_batch_item (include/psm.h)."""
    _fields_ = [("survey", Survey), ("data", C.c_void_p), ("vel", C.c_void_p),
                ("out", C.c_void_p), ("alpha", C.c_float)]


class VelocityPoint(C.Structure):
    """psm_velocity_point (include/psm.h)."""
    _fields_ = [("inline_label", C.c_int32), ("crossline_label", C.c_int32),
                ("time", C.c_double), ("velocity", C.c_float)]


class Timing(C.Structure):
    """psm_timing (include/psm.h)."""
    _fields_ = [("pack_ms", C.c_float), ("migrate_ms", C.c_float)]


class PsmError(RuntimeError):
    pass


_lib = None
_lock = threading.Lock()

# (name, restype, argtypes) of every exported symbol
SIGNATURES = [
    ("psm_version", C.c_char_p, []),
    ("psm_create", C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    ("psm_destroy", None, [C.c_void_p]),
    ("psm_last_error", C.c_char_p, [C.c_void_p]),
    ("psm_set_stream", C.c_int, [C.c_void_p, C.c_void_p]),
    ("psm_get_stream", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("psm_synchronize", C.c_int, [C.c_void_p]),
    ("psm_kirchhoff_inline", C.c_int, [
        C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_void_p, C.c_void_p,
        C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int32, C.c_int32,
        C.c_float, C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_int]),
    ("psm_survey_plan", C.c_int, [C.POINTER(Survey), C.c_void_p, C.c_int, C.POINTER(Plan)]),
    ("psm_plan_free", None, [C.POINTER(Plan)]),
    ("psm_migrate_volume", C.c_int, [C.c_void_p, C.POINTER(Survey), C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_int]),
    ("psm_migrate_volume_noalloc", C.c_int, [C.c_void_p, C.POINTER(Survey), C.c_void_p,
                                             C.c_void_p, C.c_void_p]),
    ("psm_count_pairs", C.c_int, [C.c_void_p, C.POINTER(Survey), C.c_void_p, C.c_int,
                                  C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p,
                                  C.c_void_p]),
    ("psm_demigrate_volume", C.c_int, [C.c_void_p, C.POINTER(Survey), C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_int]),
    ("psm_set_velocity_scale", C.c_int, [C.c_void_p, C.c_float, C.c_float]),
    ("psm_migrate_volume_dalpha", C.c_int, [C.c_void_p, C.POINTER(Survey), C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.c_int]),
    ("psm_dot_f64", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int,
                              C.POINTER(C.c_double)]),
    ("psm_batch_create", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                   C.POINTER(C.c_void_p)]),
    ("psm_batch_set_shape", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    ("psm_batch_autotune", C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_float)]),
    ("psm_batch_get_shape", C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("psm_batch_run", C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    ("psm_batch_destroy", None, [C.c_void_p]),
    ("psm_table_velocity", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t,
                                     C.c_void_p, C.c_size_t, C.c_size_t, C.c_double, C.c_void_p,
                                     C.c_int]),
    ("psm_alloc_host", C.c_int, [C.c_size_t, C.POINTER(C.c_void_p)]),
    ("psm_free_host", None, [C.c_void_p]),
    ("psm_set_storage", C.c_int, [C.c_void_p, C.c_int]),
    ("psm_set_tuning", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    ("psm_get_tuning", C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("psm_set_staging", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    ("psm_get_staging", C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("psm_enable_timing", C.c_int, [C.c_void_p, C.c_int]),
    ("psm_last_timing", C.c_int, [C.c_void_p, C.POINTER(Timing)]),
]


def _share_torch_hip_runtime():
    """PyTorch ships its own libamdhip64.so with the same soname
    (libamdhip64.so.7) as /opt/rocm's.  Loading torch first makes the dynamic
    linker resolve this library's HIP dependency to that already-loaded
    runtime, so the process holds one HIP runtime and torch tensors, streams
    and this library's calls share it.  (Loaded the other way round, torch
    would load a second runtime that cannot see the GPU.)"""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load():
    """Loads libpsm_hip.so and declares its signatures (no device is touched)."""
    global _lib
    with _lock:
        if _lib is None:
            _share_torch_hip_runtime()
            if not os.path.exists(LIB_PATH):
                raise PsmError(
                    f"libpsm_hip.so is not built ({LIB_PATH}); run __graft_entry__.build()")
            L = C.CDLL(LIB_PATH)
            for name, res, args in SIGNATURES:
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


class Context:
    """One psm_ctx (a HIP stream plus device buffers) on one device."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.psm_create(C.byref(h), int(device))
        if rc != PSM_OK:
            raise PsmError(f"psm_create(device={device}) failed with code {rc} "
                           "(no usable gfx950 device?)")
        self.h = h
        self.device = device

    def check(self, rc, what):
        if rc != PSM_OK:
            msg = self.lib.psm_last_error(self.h)
            raise PsmError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def set_stream(self, stream_ptr: int | None):
        """Runs later calls on this HIP stream (0/None = the null stream)."""
        self.check(self.lib.psm_set_stream(self.h, C.c_void_p(stream_ptr or 0)), "psm_set_stream")

    def stream_ptr(self) -> int:
        """The HIP stream (as an integer handle, 0 = null stream) this
        context's calls enqueue on."""
        s = C.c_void_p()
        self.check(self.lib.psm_get_stream(self.h, C.byref(s)), "psm_get_stream")
        return s.value or 0

    def synchronize(self):
        self.check(self.lib.psm_synchronize(self.h), "psm_synchronize")

    def set_storage(self, storage: str = "f32"):
        """Trace storage in HBM: 'f32', 'f16' or 'bf16' (per-trace scaled)."""
        if storage not in STORAGE:
            raise PsmError(f"unknown storage {storage!r}")
        self.check(self.lib.psm_set_storage(self.h, STORAGE[storage]), "psm_set_storage")

    def set_velocity_scale(self, alpha: float = 1.0, alpha_mask: float | None = None):
        """Travel times and weights use alpha * Vrms, the aperture mask
        alpha_mask * Vrms (default: alpha)."""
        am = alpha if alpha_mask is None else alpha_mask
        self.check(self.lib.psm_set_velocity_scale(self.h, float(alpha), float(am)),
                   "psm_set_velocity_scale")

    def set_tuning(self, waves_per_block: int = 4, rows_per_pass: int = -1):
        self.check(self.lib.psm_set_tuning(self.h, int(waves_per_block), int(rows_per_pass)),
                   "psm_set_tuning")

    def get_tuning(self):
        w, r = C.c_int(), C.c_int()
        self.check(self.lib.psm_get_tuning(self.h, C.byref(w), C.byref(r)), "psm_get_tuning")
        return w.value, r.value


    def set_staging(self, on=None, steps: int = 4):
        """Forward kernel form: aperture windows staged in LDS (True), gathered
        through the vector memory path (False), or chosen by trace storage
        (None, the default); steps = crossline offsets per staged group (2 or 4)."""
        mode = -1 if on is None else (1 if on else 0)
        self.check(self.lib.psm_set_staging(self.h, mode, int(steps)), "psm_set_staging")

    def get_staging(self):
        on, st = C.c_int(), C.c_int()
        self.check(self.lib.psm_get_staging(self.h, C.byref(on), C.byref(st)), "psm_get_staging")
        return (None if on.value < 0 else bool(on.value)), st.value
    def enable_timing(self, on: bool = True):
        self.check(self.lib.psm_enable_timing(self.h, 1 if on else 0), "psm_enable_timing")

    def last_timing(self):
        """(pack_ms, migrate_ms) of the last volume call (HIP events)."""
        t = Timing()
        self.check(self.lib.psm_last_timing(self.h, C.byref(t)), "psm_last_timing")
        return t.pack_ms, t.migrate_ms

    def close(self):
        if getattr(self, "h", None):
            self.lib.psm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_ctx = {}


def context(device: int = 0) -> Context:
    """Process-wide default context per device."""
    with _lock:
        c = _ctx.get(device)
    if c is None:
        c = Context(device)
        with _lock:
            _ctx[device] = c
    return c
