"""ctypes bindings to the checker libraries (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It loads

  oracle/libpsm_oracle.so   our C restatement of the reference path
  oracle/_ref/libpsm_ref.so the reference's own kernel, compiled from its
                            sources by oracle/Makefile (present only where it
                            was built)

and exposes numpy-level helpers.  See psm_oracle.c for what each function
restates (reference file:line).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libpsm_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libpsm_ref.so")
REF_CLI = os.path.join(HERE, "_ref", "poststackmig")

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class Volume(C.Structure):
    _fields_ = [
        ("n_il", C.c_size_t), ("n_xl", C.c_size_t), ("n_t", C.c_size_t),
        ("dt", C.c_double), ("il_step", C.c_double), ("xl_step", C.c_double),
        ("angle_deg", C.c_double),
        ("pad_il", C.c_int), ("pad_xl", C.c_int), ("amp", C.c_int),
    ]


def _f(a):
    return a.ctypes.data_as(_fp)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return None if a is None else a.ctypes.data_as(_ip)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle library not built: {ORACLE_SO} (run __graft_entry__.build())")
        L = C.CDLL(ORACLE_SO)
        L.psmo_kernel_inline_f32.argtypes = [
            _fp, C.c_size_t, C.c_size_t, C.c_size_t, _fp, _fp, C.c_size_t, _ip,
            C.c_size_t, C.c_int32, C.c_int32, C.c_float, _fp, C.c_float, C.c_float,
            C.c_float, C.c_int]
        L.psmo_kernel_inline_f64.argtypes = [
            _fp, C.c_size_t, C.c_size_t, C.c_size_t, _fp, _dp, C.c_size_t, _ip,
            C.c_size_t, C.c_int32, C.c_int32, C.c_double, _fp, C.c_float, C.c_double,
            C.c_double, C.c_int]
        L.psmo_estimate_max_velocity.argtypes = [_fp, C.c_size_t, C.c_size_t, C.c_size_t]
        L.psmo_estimate_max_velocity.restype = C.c_float
        L.psmo_aperture.argtypes = [C.c_double, C.c_float, C.c_size_t, C.c_double,
                                    C.c_double, _fp, C.POINTER(C.c_int)]
        L.psmo_output_labels.argtypes = [_ip, C.c_size_t, C.c_int, _ip]
        L.psmo_output_labels.restype = C.c_size_t
        L.psmo_map_output_inline.argtypes = [_ip, C.c_size_t, C.c_int32]
        L.psmo_map_output_inline.restype = C.c_size_t
        for name, op in (("psmo_migrate_volume_f32", _fp), ("psmo_migrate_volume_f64", _dp)):
            getattr(L, name).argtypes = [_fp, _fp, C.POINTER(Volume), _ip, _ip, op,
                                         C.c_size_t, C.c_size_t, _ip, C.c_size_t, C.c_int]
        L.psmo_adjoint_volume_f64.argtypes = [_dp, _fp, C.POINTER(Volume), _dp,
                                              C.c_size_t, C.c_size_t]
        L.psmo_migrate_sample.argtypes = [_fp, _fp, C.POINTER(Volume), _ip, C.c_size_t, C.c_int,
                                          C.c_size_t, _fp, C.c_int]
        L.psmo_count_pairs.argtypes = [_fp, C.POINTER(Volume), _ip, C.c_size_t, _dp, _dp, C.c_int]
        _lib = L
    return _lib


def have_ref() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not have_ref():
            raise RuntimeError(f"reference library not built: {REF_SO}")
        L = C.CDLL(REF_SO)
        L.psmr_kernel_inline.argtypes = [
            _fp, C.c_size_t, C.c_size_t, C.c_size_t, _fp, _fp, C.c_size_t, _ip,
            C.c_size_t, C.c_int32, C.c_int32, C.c_float, _fp, C.c_float, C.c_float,
            C.c_float, C.c_int]
        L.psmr_migrate_sample.argtypes = [
            _fp, _fp, C.c_size_t, C.c_size_t, C.c_size_t, C.c_double, C.c_double,
            C.c_double, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, _ip, _ip, C.c_size_t,
            C.c_int, C.c_size_t, _fp, C.c_int]
        L.psmr_migrate_volume.argtypes = [
            _fp, _fp, C.c_size_t, C.c_size_t, C.c_size_t, C.c_double, C.c_double,
            C.c_double, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, _ip, _fp,
            C.c_size_t, C.c_size_t, _ip, C.c_size_t, C.c_int]
        _ref = L
    return _ref


# ---------------------------------------------------------------- per call
def kernel_inline(buf, vel, il_pos, cur_il_pos, cur_xl_start, dt, t0, tan2, dx, dy,
                  amp, n_xl=None, impl="oracle"):
    """One reference kernel call.  buf [n_buf_il][n_buf_xl][n_t] f32,
    vel [n_buf_xl][n_t] f32 -> out [n_xl][n_t] (f32, or f64 for impl='f64')."""
    buf = np.ascontiguousarray(buf, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    nb, K, nt = buf.shape
    n_xl = K if n_xl is None else n_xl
    il_pos = np.ascontiguousarray(il_pos, np.int32)
    t0 = np.ascontiguousarray(t0, np.float32)
    if impl == "f64":
        out = np.zeros((n_xl, nt), np.float64)
        lib().psmo_kernel_inline_f64(_f(buf), nb, K, nt, _f(vel), _d(out), n_xl, _i(il_pos),
                                     il_pos.size, cur_il_pos, cur_xl_start, float(dt), _f(t0),
                                     float(tan2), float(dx), float(dy), int(amp))
        return out
    out = np.zeros((n_xl, nt), np.float32)
    fn = lib().psmo_kernel_inline_f32 if impl == "oracle" else ref().psmr_kernel_inline
    fn(_f(buf), nb, K, nt, _f(vel), _f(out), n_xl, _i(il_pos), il_pos.size, cur_il_pos,
       cur_xl_start, float(dt), _f(t0), float(tan2), float(dx), float(dy), int(amp))
    return out


# ---------------------------------------------------------------- geometry
def estimate_max_velocity(vel):
    vel = np.ascontiguousarray(vel, np.float32)
    return float(lib().psmo_estimate_max_velocity(_f(vel), *vel.shape))


def aperture(angle_deg, vmax, n_t, dt, il_step):
    t2 = C.c_float()
    ap = C.c_int()
    lib().psmo_aperture(angle_deg, vmax, n_t, dt, il_step, C.byref(t2), C.byref(ap))
    return t2.value, ap.value


def output_labels(labels, pad):
    labels = np.ascontiguousarray(labels, np.int32)
    n = lib().psmo_output_labels(_i(labels), labels.size, pad, None)
    out = np.zeros(max(n, 1), np.int32)
    lib().psmo_output_labels(_i(labels), labels.size, pad, _i(out))
    return out[:n]


def map_output_inline(labels, out_label):
    labels = np.ascontiguousarray(labels, np.int32)
    return int(lib().psmo_map_output_inline(_i(labels), labels.size, int(out_label)))


def _vol(data, dt, dx, dy, pad_il, pad_xl, angle, amp):
    n_il, n_xl, n_t = data.shape
    return Volume(n_il, n_xl, n_t, dt, dx, dy, angle, pad_il, pad_xl, int(amp))


def out_shape(data_shape, pad_il, pad_xl, il_labels=None):
    """Output inlines and crosslines as main.cpp:140-171 labels them (a single
    inline or crossline label gives no label step, hence no padding on that
    axis; the kernel's buffer is padded all the same, main.cpp:296-307)."""
    n_il, n_xl, n_t = data_shape
    ill = np.arange(1, n_il + 1, dtype=np.int32) if il_labels is None else il_labels
    xll = np.arange(1, n_xl + 1, dtype=np.int32)
    return output_labels(ill, pad_il).size, output_labels(xll, pad_xl).size, n_t


# ---------------------------------------------------------------- volume
def migrate_volume(data, vel, dt, dx, dy, pad_il=0, pad_xl=0, angle=30.0, amp=True,
                   rows=None, nthreads=0, precision="f32", il_labels=None):
    """main.cpp's whole loop on an in-memory cube (oracle restatement)."""
    data = np.ascontiguousarray(data, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    g = _vol(data, dt, dx, dy, pad_il, pad_xl, angle, amp)
    shp = out_shape(data.shape, pad_il, pad_xl, il_labels)
    out = np.zeros(shp, np.float32 if precision == "f32" else np.float64)
    rws = None if rows is None else np.ascontiguousarray(rows, np.int32)
    ill = None if il_labels is None else np.ascontiguousarray(il_labels, np.int32)
    fn = lib().psmo_migrate_volume_f32 if precision == "f32" else lib().psmo_migrate_volume_f64
    rc = fn(_f(data), _f(vel), C.byref(g), _i(ill), None,
            _f(out) if precision == "f32" else _d(out), shp[0], shp[1], _i(rws),
            0 if rws is None else rws.size, nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle migrate_volume failed ({rc})")
    return out


def ref_migrate_volume(data, vel, dt, dx, dy, pad_il=0, pad_xl=0, angle=30.0, amp=True,
                       rows=None, nthreads=0):
    """The reference's compiled kernel driven by main.cpp's loop (oracle/_ref)."""
    data = np.ascontiguousarray(data, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    n_il, n_xl, n_t = data.shape
    vmax = estimate_max_velocity(vel)
    t2, ap = aperture(angle, vmax, n_t, dt, dx)
    ill = np.arange(1, n_il + 1, dtype=np.int32)
    oll = output_labels(ill, pad_il)
    c_of = np.array([map_output_inline(ill, l) for l in oll], np.int32)
    shp = out_shape(data.shape, pad_il, pad_xl)
    out = np.zeros(shp, np.float32)
    rws = None if rows is None else np.ascontiguousarray(rows, np.int32)
    ref().psmr_migrate_volume(_f(data), _f(vel), n_il, n_xl, n_t, dt, dx, dy, pad_il, pad_xl,
                              t2, ap, int(amp), _i(c_of), _f(out), shp[0], shp[1], _i(rws),
                              0 if rws is None else rws.size, nthreads)
    return out


def ref_migrate_sample(data, vel, dt, dx, dy, pad_il, pad_xl, angle, amp, rows, xl0, m,
                       nthreads=0):
    """Reference kernel on a bounded sample: output inlines `rows`, output
    crosslines [xl0, xl0+m) of each (oracle/_ref).  Returns [len(rows)][m][n_t]."""
    data = np.ascontiguousarray(data, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    n_il, n_xl, n_t = data.shape
    vmax = estimate_max_velocity(vel)
    t2, ap = aperture(angle, vmax, n_t, dt, dx)
    ill = np.arange(1, n_il + 1, dtype=np.int32)
    oll = output_labels(ill, pad_il)
    c_of = np.array([map_output_inline(ill, l) for l in oll], np.int32)
    rws = np.ascontiguousarray(rows, np.int32)
    out = np.zeros((rws.size, m, n_t), np.float32)
    ref().psmr_migrate_sample(_f(data), _f(vel), n_il, n_xl, n_t, dt, dx, dy, pad_il, pad_xl,
                              t2, ap, int(amp), _i(c_of), _i(rws), rws.size, int(xl0), int(m),
                              _f(out), nthreads)
    return out


def migrate_sample(data, vel, dt, dx, dy, pad_il, pad_xl, angle, amp, rows, xl0, m,
                   nthreads=0):
    """Oracle restatement on the same bounded sample as ref_migrate_sample."""
    data = np.ascontiguousarray(data, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    g = _vol(data, dt, dx, dy, pad_il, pad_xl, angle, amp)
    rws = np.ascontiguousarray(rows, np.int32)
    out = np.zeros((rws.size, m, data.shape[2]), np.float32)
    lib().psmo_migrate_sample(_f(data), _f(vel), C.byref(g), _i(rws), rws.size, int(xl0), int(m),
                              _f(out), nthreads)
    return out


def adjoint_volume_f64(image, vel, dt, dx, dy, pad_il=0, pad_xl=0, angle=30.0, amp=True):
    image = np.ascontiguousarray(image, np.float64)
    vel = np.ascontiguousarray(vel, np.float32)
    n_il, n_xl, n_t = vel.shape
    g = Volume(n_il, n_xl, n_t, dt, dx, dy, angle, pad_il, pad_xl, int(amp))
    out = np.zeros(vel.shape, np.float64)
    rc = lib().psmo_adjoint_volume_f64(_d(image), _f(vel), C.byref(g), _d(out),
                                       image.shape[0], image.shape[1])
    if rc != 0:
        raise RuntimeError("oracle adjoint failed")
    return out


def count_pairs(vel, dt, dx, dy, pad_il=0, pad_xl=0, angle=30.0, rows=None, nthreads=0):
    vel = np.ascontiguousarray(vel, np.float32)
    n_il, n_xl, n_t = vel.shape
    g = Volume(n_il, n_xl, n_t, dt, dx, dy, angle, pad_il, pad_xl, 1)
    rws = None if rows is None else np.ascontiguousarray(rows, np.int32)
    pap = C.c_double()
    pg = C.c_double()
    lib().psmo_count_pairs(_f(vel), C.byref(g), _i(rws), 0 if rws is None else rws.size,
                           C.byref(pap), C.byref(pg), nthreads)
    return pap.value, pg.value


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))
