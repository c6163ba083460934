"""Seeded synthetic post-stack cubes for the migration workloads.

All cubes are float32 [n_il][n_xl][n_t] on a regular grid (labels 1..n),
dt = 2 ms, 25 m bins unless stated.  Velocity cubes are Vrms(il, xl, t) on the
same grid, as the reference's SEG-Y velocity provider reads them on its
exact-grid path (velocity_reader.cpp:216-231).

Workloads
  point_diffractor   one Ricker diffraction hyperbola at the cube centre plus
                     seeded Gaussian noise (101 x 101 x 501, V = 2000 m/s,
                     25 Hz, 10% noise is the plain-method CPU case)
  production         401 x 401 x 1001, Vrms = 1500 + 1200 t
                     + 150 sin(2 pi il/400) cos(2 pi xl/400), 64 seeded point
                     diffractors + 3 planar dipping reflectors, 30 Hz Ricker
  spikes             sparse reflectivity (0.5 % density, N(0,1) amplitudes)
"""
from __future__ import annotations

import math

import numpy as np


def ricker(t, f):
    a = (math.pi * f * t) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


def _add_event(cube, tau, f, dt, amp=1.0, half_width=None):
    """Adds amp * ricker(t - tau[il, xl]) to every trace, evaluated only on a
    window of samples around tau (the wavelet is ~0 beyond 1.5/f).  Within
    one event every (trace, sample) index is distinct, so a plain fancy-index
    add is exact."""
    n_il, n_xl, n_t = cube.shape
    hw = half_width if half_width is not None else int(math.ceil(1.5 / f / dt)) + 1
    i0 = np.floor(tau / dt).astype(np.int64)
    offs = np.arange(-hw, hw + 2, dtype=np.int64)
    idx = i0[..., None] + offs
    valid = (idx >= 0) & (idx < n_t)
    w = (amp * ricker(idx * dt - tau[..., None], f)).astype(np.float32)
    base = (np.arange(n_il, dtype=np.int64)[:, None] * n_xl + np.arange(n_xl)[None, :]) * n_t
    fidx = (base[..., None] + idx)[valid]
    flat = cube.reshape(-1)
    flat[fidx] += w[valid]


def constant_velocity(shape, v):
    return np.full(shape, np.float32(v), dtype=np.float32)


def point_diffractor(n_il=101, n_xl=101, n_t=501, dt=0.002, dx=25.0, dy=25.0,
                     v=2000.0, f=25.0, noise=0.1, seed=1, t_apex=None):
    """Diffraction hyperbola tau = sqrt(t_a^2 + 4 r^2 / v^2) with its apex at the
    cube centre (il, xl) = (n_il//2, n_xl//2), t_a = centre time by default,
    peak amplitude 1, plus N(0, noise^2) from `seed`.  Returns (data, vel)."""
    ta = (n_t - 1) * dt / 2.0 if t_apex is None else t_apex
    il = (np.arange(n_il) - n_il // 2) * dx
    xl = (np.arange(n_xl) - n_xl // 2) * dy
    r2 = il[:, None] ** 2 + xl[None, :] ** 2
    tau = np.sqrt(ta * ta + 4.0 * r2 / (v * v))
    cube = np.zeros((n_il, n_xl, n_t), np.float32)
    _add_event(cube, tau, f, dt)
    if noise > 0:
        rng = np.random.default_rng(seed)
        cube += rng.normal(0.0, noise, cube.shape).astype(np.float32)
    return cube, constant_velocity(cube.shape, v)


def production_velocity(n_il=401, n_xl=401, n_t=1001, dt=0.002):
    t = (np.arange(n_t) * dt).astype(np.float64)
    il = np.arange(n_il, dtype=np.float64)
    xl = np.arange(n_xl, dtype=np.float64)
    lat = 150.0 * np.sin(2 * np.pi * il / 400.0)[:, None] * np.cos(2 * np.pi * xl / 400.0)[None, :]
    v = np.empty((n_il, n_xl, n_t), np.float32)
    base = (1500.0 + 1200.0 * t).astype(np.float32)
    v[...] = base[None, None, :]
    v += lat.astype(np.float32)[..., None]
    return v


def production(n_il=401, n_xl=401, n_t=1001, dt=0.002, dx=25.0, dy=25.0, f=30.0,
               n_diffractors=64, seed=2):
    """64 seeded point diffractors (hyperbolae with the Vrms at the apex) plus 3
    planar dipping reflectors (dips 0..20 degrees, time dip from the local
    Vrms), all 30 Hz Ricker.  Returns (data, vel)."""
    rng = np.random.default_rng(seed)
    vel = production_velocity(n_il, n_xl, n_t, dt)
    cube = np.zeros((n_il, n_xl, n_t), np.float32)
    X = (np.arange(n_il) * dx)[:, None]
    Y = (np.arange(n_xl) * dy)[None, :]
    for _ in range(n_diffractors):
        ia = rng.integers(0, n_il)
        xa = rng.integers(0, n_xl)
        ta = rng.uniform(min(0.2, 0.3 * (n_t - 1) * dt), (n_t - 1) * dt * 0.9)
        va = float(vel[ia, xa, min(int(ta / dt), n_t - 1)])
        r2 = (X - ia * dx) ** 2 + (Y - xa * dy) ** 2
        tau = np.sqrt(ta * ta + 4.0 * r2 / (va * va))
        _add_event(cube, tau, f, dt, amp=rng.uniform(0.5, 1.0))
    for k in range(3):
        dip = np.deg2rad(rng.uniform(0.0, 20.0))
        az = rng.uniform(0.0, 2 * np.pi)
        t_c = (k + 1) * (n_t - 1) * dt / 4.0
        vc = float(vel[n_il // 2, n_xl // 2, min(int(t_c / dt), n_t - 1)])
        # zero-offset two-way time of a plane dipping at `dip` along azimuth az
        p = 2.0 * np.sin(dip) / vc
        tau = t_c + p * (np.cos(az) * (X - n_il * dx / 2) + np.sin(az) * (Y - n_xl * dy / 2))
        _add_event(cube, tau, f, dt, amp=0.8)
    return cube, vel


def production_torch(device, n_il=401, n_xl=401, n_t=1001, dt=0.002, dx=25.0, dy=25.0,
                     f=30.0, n_diffractors=64, seed=2):
    """production() evaluated with torch on `device` (same seeded event list;
    the wavelet is summed in float32 on the device).  Returns (data, vel) as
    device tensors."""
    import torch

    rng = np.random.default_rng(seed)
    vel = torch.from_numpy(production_velocity(n_il, n_xl, n_t, dt)).to(device)
    cube = torch.zeros((n_il, n_xl, n_t), dtype=torch.float32, device=device)
    X = (torch.arange(n_il, device=device, dtype=torch.float64) * dx)[:, None]
    Y = (torch.arange(n_xl, device=device, dtype=torch.float64) * dy)[None, :]
    hw = int(math.ceil(1.5 / f / dt)) + 1
    offs = torch.arange(-hw, hw + 2, device=device, dtype=torch.int64)
    base = ((torch.arange(n_il, device=device)[:, None] * n_xl
             + torch.arange(n_xl, device=device)[None, :]) * n_t)
    flat = cube.view(-1)

    def add(tau, amp):
        i0 = torch.floor(tau / dt).to(torch.int64)
        idx = i0[..., None] + offs
        valid = (idx >= 0) & (idx < n_t)
        a = (math.pi * f * (idx * dt - tau[..., None])) ** 2
        w = (amp * (1.0 - 2.0 * a) * torch.exp(-a)).to(torch.float32)
        flat.index_add_(0, (base[..., None] + idx)[valid], w[valid])

    velc = vel.cpu().numpy()
    for _ in range(n_diffractors):
        ia = rng.integers(0, n_il)
        xa = rng.integers(0, n_xl)
        ta = rng.uniform(min(0.2, 0.3 * (n_t - 1) * dt), (n_t - 1) * dt * 0.9)
        va = float(velc[ia, xa, min(int(ta / dt), n_t - 1)])
        r2 = (X - ia * dx) ** 2 + (Y - xa * dy) ** 2
        add(torch.sqrt(ta * ta + 4.0 * r2 / (va * va)), rng.uniform(0.5, 1.0))
    for k in range(3):
        dip = np.deg2rad(rng.uniform(0.0, 20.0))
        az = rng.uniform(0.0, 2 * np.pi)
        t_c = (k + 1) * (n_t - 1) * dt / 4.0
        vc = float(velc[n_il // 2, n_xl // 2, min(int(t_c / dt), n_t - 1)])
        p = 2.0 * np.sin(dip) / vc
        add(t_c + p * (np.cos(az) * (X - n_il * dx / 2) + np.sin(az) * (Y - n_xl * dy / 2)), 0.8)
    return cube, vel


def spikes(shape, density=0.005, seed=3):
    """Sparse reflectivity: a fraction `density` of samples ~ N(0, 1)."""
    rng = np.random.default_rng(seed)
    m = np.zeros(shape, np.float32)
    n = int(round(density * m.size))
    idx = rng.choice(m.size, size=n, replace=False)
    m.ravel()[idx] = rng.normal(0.0, 1.0, n).astype(np.float32)
    return m
