"""ctypes bindings of libpsm_host.so (include/psm_host.h): configuration
file, SEG-Y ingest and output, velocity providers and the whole
`poststackmig` run.

    cfg = host_api.parse_config("config.txt")
    with host_api.SegyFile(path) as f:
        cube = f.read_cube()                  # [n_il][n_xl][n_t] float32, one pass
    vel = host_api.velocity_cube("vel.sgy", f.inline_labels, f.crossline_labels, n_t, dt)
    host_api.run("config.txt")                # == bin/poststackmig config.txt
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

from . import _lib
from ._lib import PSM_OK, PsmError

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsm_host.so")
CLI_PATH = os.path.join(HERE, "bin", "poststackmig")

_lock = threading.Lock()
_host = None


class Config(C.Structure):
    """psmh_config."""
    _fields_ = [("input_data", C.c_char * 1024), ("output_data", C.c_char * 1024),
                ("velocity", C.c_char * 1024), ("inline_step", C.c_double),
                ("crossline_step", C.c_double), ("inline_padding", C.c_int),
                ("crossline_padding", C.c_int), ("angle_aperture", C.c_double),
                ("amp_correction", C.c_int), ("n_threads", C.c_int)]

    def as_dict(self):
        return {k: (getattr(self, k).decode() if isinstance(getattr(self, k), bytes)
                    else getattr(self, k)) for k, _ in self._fields_}


SIGNATURES = [
    ("psmh_last_error", C.c_char_p, []),
    ("psmh_config_parse", C.c_int, [C.c_char_p, C.POINTER(Config)]),
    ("psmh_segy_open", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    ("psmh_segy_close", None, [C.c_void_p]),
    ("psmh_segy_info", C.c_int, [C.c_void_p, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t),
                                 C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_size_t),
                                 C.POINTER(C.c_size_t)]),
    ("psmh_segy_inline_labels", C.c_void_p, [C.c_void_p]),
    ("psmh_segy_crossline_labels", C.c_void_p, [C.c_void_p]),
    ("psmh_segy_read_cube", C.c_int, [C.c_void_p, C.c_void_p]),
    ("psmh_ibm_to_ieee", None, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("psmh_ieee_to_ibm", None, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("psmh_segy_write", C.c_int, [C.c_char_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                  C.c_size_t, C.c_void_p, C.c_size_t]),
    ("psmh_velocity_cube", C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                     C.c_size_t, C.c_size_t, C.c_double, C.c_void_p]),
    ("psmh_run", C.c_int, [C.c_char_p, C.c_int]),
]


def load():
    """Loads libpsm_host.so (and libpsm_hip.so through _lib) without touching a device."""
    global _host
    with _lock:
        if _host is None:
            _lib.load()
            if not os.path.exists(LIB_PATH):
                raise PsmError(f"libpsm_host.so is not built ({LIB_PATH}); run __graft_entry__.build()")
            L = C.CDLL(LIB_PATH)
            for name, res, args in SIGNATURES:
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _host = L
    return _host


def _check(rc, what):
    if rc != PSM_OK:
        raise PsmError(f"{what}: {load().psmh_last_error().decode()} (status {rc})")


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def parse_config(path):
    cfg = Config()
    _check(load().psmh_config_parse(os.fsencode(path), C.byref(cfg)), "config")
    return cfg.as_dict()


class SegyFile:
    """An opened SEG-Y file: metadata, sorted labels and the one-pass cube read."""

    def __init__(self, path):
        self.lib = load()
        h = C.c_void_p()
        _check(self.lib.psmh_segy_open(os.fsencode(path), C.byref(h)), f"open {path}")
        self.h = h
        nt, n, nil, nxl = C.c_size_t(), C.c_size_t(), C.c_size_t(), C.c_size_t()
        dt, fmt = C.c_int(), C.c_int()
        self.lib.psmh_segy_info(h, C.byref(n), C.byref(nt), C.byref(dt), C.byref(fmt),
                                C.byref(nil), C.byref(nxl))
        self.n_traces, self.n_t, self.dt_us, self.fmt = n.value, nt.value, dt.value, fmt.value
        self.n_il, self.n_xl = nil.value, nxl.value
        ip = C.cast(self.lib.psmh_segy_inline_labels(h), C.POINTER(C.c_int32))
        xp = C.cast(self.lib.psmh_segy_crossline_labels(h), C.POINTER(C.c_int32))
        self.inline_labels = np.ctypeslib.as_array(ip, (self.n_il,)).copy()
        self.crossline_labels = np.ctypeslib.as_array(xp, (self.n_xl,)).copy()

    def read_cube(self):
        cube = np.empty((self.n_il, self.n_xl, self.n_t), np.float32)
        _check(self.lib.psmh_segy_read_cube(self.h, _ptr(cube)), "read cube")
        return cube

    def write(self, path, out_il, out_xl, image):
        """Output file with this file's headers (psmh_segy_write)."""
        out_il = np.ascontiguousarray(out_il, np.int32)
        out_xl = np.ascontiguousarray(out_xl, np.int32)
        image = np.ascontiguousarray(image, np.float32)
        assert image.shape == (out_il.size, out_xl.size, self.n_t)
        _check(self.lib.psmh_segy_write(os.fsencode(path), self.h, _ptr(out_il), out_il.size,
                                        _ptr(out_xl), out_xl.size, _ptr(image), self.n_t),
               f"write {path}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.psmh_segy_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()


def ibm_to_ieee(u):
    u = np.ascontiguousarray(u, np.uint32)
    out = np.empty(u.shape, np.float32)
    load().psmh_ibm_to_ieee(_ptr(u), _ptr(out), u.size)
    return out


def ieee_to_ibm(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty(x.shape, np.uint32)
    load().psmh_ieee_to_ibm(_ptr(x), _ptr(out), x.size)
    return out


def velocity_cube(param, il_labels, xl_labels, n_t, dt, ctx: _lib.Context | None = None):
    """Vrms cube [n_il][n_xl][n_t] on the data grid from a velocity parameter
    (number, SEG-Y path or text table; the table needs a GPU context)."""
    il = np.ascontiguousarray(il_labels, np.int32)
    xl = np.ascontiguousarray(xl_labels, np.int32)
    out = np.empty((il.size, xl.size, int(n_t)), np.float32)
    h = ctx.h if ctx is not None else None
    _check(load().psmh_velocity_cube(h, os.fsencode(str(param)), _ptr(il), il.size, _ptr(xl),
                                     xl.size, int(n_t), float(dt), _ptr(out)), "velocity")
    return out


def table_velocity(points, il_labels, xl_labels, n_t, dt, ctx: _lib.Context | None = None):
    """psm_table_velocity on the GPU: points = rows (il, xl, time, vel)."""
    ctx = ctx or _lib.context()
    pts = (_lib.VelocityPoint * len(points))()
    for k, (a, b, t, v) in enumerate(points):
        pts[k] = _lib.VelocityPoint(int(a), int(b), float(t), float(v))
    il = np.ascontiguousarray(il_labels, np.int32)
    xl = np.ascontiguousarray(xl_labels, np.int32)
    out = np.empty((il.size, xl.size, int(n_t)), np.float32)
    ctx.check(ctx.lib.psm_table_velocity(ctx.h, pts, len(points), _ptr(il), il.size, _ptr(xl),
                                         xl.size, int(n_t), float(dt), _ptr(out), _lib.PSM_HOST),
              "psm_table_velocity")
    return out


def run(config_path, device: int = 0) -> int:
    """The whole poststackmig run in this process; returns the exit status."""
    return load().psmh_run(os.fsencode(config_path), int(device))
