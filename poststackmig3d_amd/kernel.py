"""Python host mirror of the reference's migration-kernel interface.

  kirchhoff_kernel_3d_with_padding   MigrationKernel::kirchhoffKernel3DWithPadding
                                     (include/migration_kernel.h:16-31): same
                                     argument names and order; output_slice is
                                     a [n_xl][n_t] float32 array written in
                                     place (the reference's
                                     vector<vector<float>>&)
  migrate_volume                     main.cpp:254-383's loop over output
                                     inlines on a whole cube
  survey_plan                        the geometry of main.cpp:140-217
  count_pairs                        work-unit pair counts

Every call goes through libpsm_hip.so (include/psm.h) to the gfx950 kernels.
Arrays may be numpy (host) or torch CUDA tensors (device, resident in HBM);
torch is only used for device memory and streams.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import PSM_DEVICE, PSM_HOST, PsmError, Survey, Plan


def _host_f32(a, name):
    arr = np.ascontiguousarray(a, dtype=np.float32)
    if arr.dtype != np.float32:
        raise PsmError(f"{name} must be float32")
    return arr


def kirchhoff_kernel_3d_with_padding(input_buffer, n_buf_il, n_buf_xl, n_t, velocity_slice,
                                     output_slice, il_pos_indices, xl_pos_indices,
                                     current_il_pos, current_xl_start, dt, t0_times,
                                     tan_theta_sq, inline_bin_size, crossline_bin_size,
                                     amp_correction, ctx: _lib.Context | None = None):
    """One output inline, as MigrationKernel::kirchhoffKernel3DWithPadding.

    input_buffer   float32, n_buf_il*n_buf_xl*n_t values, [il][xl][t]
    velocity_slice float32, n_buf_xl*n_t values, [xl][t]
    output_slice   float32 array [n_xl][n_t], fully overwritten
    """
    ctx = ctx or _lib.context()
    n_buf_il, n_buf_xl, n_t = int(n_buf_il), int(n_buf_xl), int(n_t)
    if not isinstance(output_slice, np.ndarray) or output_slice.dtype != np.float32 \
            or output_slice.ndim != 2 or not output_slice.flags.c_contiguous:
        raise PsmError("output_slice must be a C-contiguous float32 [n_xl][n_t] array")
    n_xl = output_slice.shape[0]
    if n_xl and output_slice.shape[1] != n_t:
        raise PsmError(f"output_slice has {output_slice.shape[1]} samples, expected {n_t}")
    buf = _host_f32(input_buffer, "input_buffer").reshape(-1)
    vel = _host_f32(velocity_slice, "velocity_slice").reshape(-1)
    if buf.size < n_buf_il * n_buf_xl * n_t:
        raise PsmError("input_buffer is smaller than n_buf_il*n_buf_xl*n_t")
    if vel.size < n_buf_xl * n_t:
        raise PsmError("velocity_slice is smaller than n_buf_xl*n_t")
    ilp = np.ascontiguousarray(il_pos_indices, dtype=np.int32)
    xlp = None if xl_pos_indices is None else np.ascontiguousarray(xl_pos_indices, dtype=np.int32)
    t0 = _host_f32(t0_times, "t0_times")
    if t0.size < n_t:
        raise PsmError("t0_times is shorter than n_t")
    rc = ctx.lib.psm_kirchhoff_inline(
        ctx.h, buf.ctypes.data, n_buf_il, n_buf_xl, n_t, vel.ctypes.data,
        output_slice.ctypes.data, n_xl, ilp.ctypes.data, ilp.size,
        None if xlp is None else xlp.ctypes.data, 0 if xlp is None else xlp.size,
        int(current_il_pos), int(current_xl_start), float(dt), t0.ctypes.data,
        float(tan_theta_sq), float(inline_bin_size), float(crossline_bin_size),
        1 if amp_correction else 0)
    ctx.check(rc, "psm_kirchhoff_inline")


@dataclass
class SurveyParams:
    """The Config fields (config_parser.h:7-20) and grid that reach the kernel."""
    n_il: int
    n_xl: int
    n_t: int
    dt: float = 0.002
    inline_step: float = 25.0
    crossline_step: float = 25.0
    inline_padding: int = 0
    crossline_padding: int = 0
    angle_aperture: float = 30.0
    amp_correction: bool = True
    inline_labels: np.ndarray | None = None
    crossline_labels: np.ndarray | None = None
    max_velocity: float = 0.0
    _keep: list = field(default_factory=list, repr=False)

    def to_c(self) -> Survey:
        s = Survey()
        s.n_il, s.n_xl, s.n_t = self.n_il, self.n_xl, self.n_t
        s.dt, s.inline_step, s.crossline_step = self.dt, self.inline_step, self.crossline_step
        s.inline_padding, s.crossline_padding = self.inline_padding, self.crossline_padding
        s.angle_aperture = self.angle_aperture
        s.amp_correction = 1 if self.amp_correction else 0
        self._keep = []
        for name in ("inline_labels", "crossline_labels"):
            lab = getattr(self, name)
            if lab is None:
                setattr(s, name, None)
            else:
                arr = np.ascontiguousarray(lab, dtype=np.int32)
                self._keep.append(arr)
                setattr(s, name, arr.ctypes.data_as(C.POINTER(C.c_int32)))
        s.max_velocity = float(self.max_velocity)
        return s

    @classmethod
    def for_cube(cls, shape, **kw):
        n_il, n_xl, n_t = shape
        return cls(n_il=n_il, n_xl=n_xl, n_t=n_t, **kw)


def _is_torch(x):
    return type(x).__module__.startswith("torch")


class _TorchOrder:
    """Orders a context's device work against torch's current stream around
    one call on CUDA tensors: the context's stream first waits for what torch
    has queued (the inputs may still be in flight there), and torch's stream
    then waits for the call's work (its outputs), so neither side reads
    unfinished memory whatever streams the two use.  Nothing is done for host
    arrays (those calls are synchronous) or when both use the same stream."""

    def __init__(self, ctx, *arrays):
        self.ctx = ctx
        self.dev = next((a.device for a in arrays if _is_torch(a) and a.is_cuda), None)
        self.ext = None

    def __enter__(self):
        if self.dev is None:
            return self
        import torch
        self.cur = torch.cuda.current_stream(self.dev)
        sp = self.ctx.stream_ptr()
        if sp != self.cur.cuda_stream:
            self.ext = torch.cuda.ExternalStream(sp, device=self.dev)
            self.ext.wait_stream(self.cur)
        return self

    def __exit__(self, *exc):
        if self.ext is not None:
            self.cur.wait_stream(self.ext)
        return False


def _ptr_where(x, name):
    """(pointer, where) for a numpy array or torch tensor."""
    if _is_torch(x):
        import torch
        if x.dtype != torch.float32 or not x.is_contiguous():
            raise PsmError(f"{name} must be a contiguous float32 tensor")
        return x.data_ptr(), (PSM_DEVICE if x.is_cuda else PSM_HOST)
    if isinstance(x, np.ndarray):
        if x.dtype != np.float32 or not x.flags.c_contiguous:
            raise PsmError(f"{name} must be a C-contiguous float32 array")
        return x.ctypes.data, PSM_HOST
    raise PsmError(f"{name}: unsupported array type {type(x)}")


@dataclass
class PlanInfo:
    n_out_il: int
    n_out_xl: int
    n_t: int
    max_aperture_il: int
    tan_theta_sq: float
    max_velocity: float
    out_inline_labels: np.ndarray
    out_crossline_labels: np.ndarray
    out_to_in_il: np.ndarray
    window: np.ndarray
    t0_times: np.ndarray


def survey_plan(params: SurveyParams, vel=None) -> PlanInfo:
    """Geometry of main.cpp:140-217 for a survey (velocity probed if needed)."""
    lib = _lib.load()
    s = params.to_c()
    p = Plan()
    ptr, where = (None, PSM_HOST) if vel is None else _ptr_where(vel, "vel")
    if where == PSM_DEVICE and not params.max_velocity > 0:
        # the probe reads vel through a synchronous copy that no stream
        # orders: let torch's queued work (which may still write vel) finish
        import torch
        torch.cuda.current_stream(vel.device).synchronize()
    rc = lib.psm_survey_plan(C.byref(s), ptr, where, C.byref(p))
    if rc != _lib.PSM_OK:
        raise PsmError(f"psm_survey_plan failed ({rc})")
    try:
        n, m, nt = p.n_out_il, p.n_out_xl, p.n_t
        info = PlanInfo(
            n, m, nt, p.max_aperture_il, p.tan_theta_sq, p.max_velocity,
            np.ctypeslib.as_array(p.out_inline_labels, (n,)).copy(),
            np.ctypeslib.as_array(p.out_crossline_labels, (m,)).copy(),
            np.ctypeslib.as_array(p.out_to_in_il, (n,)).copy(),
            np.ctypeslib.as_array(p.window, (2 * n,)).reshape(n, 2).copy(),
            np.ctypeslib.as_array(p.t0_times, (nt,)).copy())
    finally:
        lib.psm_plan_free(C.byref(p))
    return info


def migrate_volume(data, vel, params: SurveyParams, out=None, ctx: _lib.Context | None = None,
                   sync: bool = True):
    """Whole-volume migration: data, vel [n_il][n_xl][n_t] float32 (numpy or
    torch CUDA) -> out [n_out_il][n_out_xl][n_t] (same kind as data)."""
    ctx = ctx or _lib.context()
    if tuple(data.shape) != (params.n_il, params.n_xl, params.n_t) or tuple(vel.shape) != tuple(data.shape):
        raise PsmError(f"data/vel shape {tuple(data.shape)}/{tuple(vel.shape)} does not match the survey "
                       f"({params.n_il}, {params.n_xl}, {params.n_t})")
    dp, dw = _ptr_where(data, "data")
    vp, vw = _ptr_where(vel, "vel")
    if dw != vw:
        raise PsmError("data and vel must both be host or both be device memory")
    if out is None:
        info = survey_plan(params, vel)
        shape = (info.n_out_il, info.n_out_xl, info.n_t)
        if _is_torch(data):
            import torch
            out = torch.empty(shape, dtype=torch.float32, device=data.device)
        else:
            out = np.empty(shape, np.float32)
    op, ow = _ptr_where(out, "out")
    if ow != dw:
        raise PsmError("out must live where data lives")
    s = params.to_c()
    with _TorchOrder(ctx, data, vel, out):
        ctx.check(ctx.lib.psm_migrate_volume(ctx.h, C.byref(s), dp, vp, op, dw),
                  "psm_migrate_volume")
        if sync and dw == PSM_DEVICE:
            ctx.synchronize()
    return out


def _volume_call(fn_name, inp, vel, params, out, out_shape, ctx, sync):
    ctx = ctx or _lib.context()
    ip, iw = _ptr_where(inp, "input")
    vp, vw = _ptr_where(vel, "vel")
    if iw != vw:
        raise PsmError("input and vel must both be host or both be device memory")
    if tuple(vel.shape) != (params.n_il, params.n_xl, params.n_t):
        raise PsmError(f"vel shape {tuple(vel.shape)} does not match the survey")
    if out is None:
        if _is_torch(inp):
            import torch
            out = torch.empty(out_shape, dtype=torch.float32, device=inp.device)
        else:
            out = np.empty(out_shape, np.float32)
    op, ow = _ptr_where(out, "out")
    if ow != iw:
        raise PsmError("out must live where the input lives")
    s = params.to_c()
    with _TorchOrder(ctx, inp, vel, out):
        ctx.check(getattr(ctx.lib, fn_name)(ctx.h, C.byref(s), ip, vp, op, iw), fn_name)
        if sync and iw == PSM_DEVICE:
            ctx.synchronize()
    return out


def demigrate_volume(image, vel, params: SurveyParams, out=None,
                     ctx: _lib.Context | None = None, sync: bool = True):
    """Adjoint of migrate_volume: image [n_out_il][n_out_xl][n_t] -> data
    [n_il][n_xl][n_t] (Kirchhoff demigration with the forward operator's
    sample index, fraction, weight and mask)."""
    info = survey_plan(params, vel)
    if tuple(image.shape) != (info.n_out_il, info.n_out_xl, info.n_t):
        raise PsmError(f"image shape {tuple(image.shape)} does not match the survey output "
                       f"({info.n_out_il}, {info.n_out_xl}, {info.n_t})")
    return _volume_call("psm_demigrate_volume", image, vel, params, out,
                        (params.n_il, params.n_xl, params.n_t), ctx, sync)


def migrate_volume_dalpha(data, vel, params: SurveyParams, out=None,
                          ctx: _lib.Context | None = None, sync: bool = True):
    """d(migrate_volume)/d(alpha) at the context's velocity scale, mask held
    at its mask scale."""
    info = survey_plan(params, vel)
    return _volume_call("psm_migrate_volume_dalpha", data, vel, params, out,
                        (info.n_out_il, info.n_out_xl, info.n_t), ctx, sync)


def dot_f64(x, y, ctx: _lib.Context | None = None) -> float:
    """sum x*y of two float32 arrays accumulated in float64 on the GPU."""
    ctx = ctx or _lib.context()
    xp, xw = _ptr_where(x, "x")
    yp, yw = _ptr_where(y, "y")
    n = int(np.prod(x.shape))
    if xw != yw or n != int(np.prod(y.shape)):
        raise PsmError("x and y must have the same size and live in the same memory")
    r = C.c_double()
    with _TorchOrder(ctx, x, y):
        ctx.check(ctx.lib.psm_dot_f64(ctx.h, xp, yp, n, xw, C.byref(r)), "psm_dot_f64")
    return r.value


def count_pairs(vel, params: SurveyParams, ctx: _lib.Context | None = None, images=False):
    """(P_ap, P_g) of the work unit for this survey, counted on the GPU.  With
    images=True also returns the per-output-sample counts (float32 arrays of
    the output shape, on the device when vel is a device tensor)."""
    ctx = ctx or _lib.context()
    vp, vw = _ptr_where(vel, "vel")
    s = params.to_c()
    pap, pg = C.c_double(), C.c_double()
    ap_img = g_img = None
    if images:
        info = survey_plan(params, vel)
        shape = (info.n_out_il, info.n_out_xl, info.n_t)
        if _is_torch(vel) and vw == PSM_DEVICE:
            import torch
            ap_img = torch.empty(shape, dtype=torch.float32, device=vel.device)
            g_img = torch.empty(shape, dtype=torch.float32, device=vel.device)
        else:
            ap_img = np.empty(shape, np.float32)
            g_img = np.empty(shape, np.float32)
    with _TorchOrder(ctx, vel, ap_img, g_img):
        ctx.check(ctx.lib.psm_count_pairs(ctx.h, C.byref(s), vp, vw, C.byref(pap), C.byref(pg),
                                          None if ap_img is None else _ptr_where(ap_img, "ap")[0],
                                          None if g_img is None else _ptr_where(g_img, "g")[0]),
                  "psm_count_pairs")
    if images:
        return pap.value, pg.value, ap_img, g_img
    return pap.value, pg.value
