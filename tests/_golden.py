"""Loads the golden fixtures of tests/golden/ (written by the reference CLI, see
tests/golden/make_golden.py) and rebuilds their inputs as cubes."""
from __future__ import annotations

import glob
import os
import tempfile

import numpy as np

from poststackmig3d_amd import segy
import velocity_oracle as V

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def parse_config(text):
    cfg = {}
    for line in str(text).splitlines():
        if "=" in line:
            k, v = line.split("=", 1)
            cfg[k.strip()] = v.strip()
    return cfg


def load(path):
    z = np.load(path, allow_pickle=False)
    d = {k: z[k] for k in z.files}
    d["name"] = os.path.splitext(os.path.basename(path))[0]
    d["cfg"] = parse_config(d["config"])
    return d


def _read_bytes(arr):
    with tempfile.NamedTemporaryFile(suffix=".sgy", delete=False) as f:
        f.write(arr.tobytes())
        name = f.name
    try:
        return segy.read(name)
    finally:
        os.unlink(name)


def inputs(fx):
    """(cube, vel_cube, il_labels, xl_labels, dt, cfg dict with numbers)."""
    seg = _read_bytes(fx["input_file"])
    cube, ill, xll = segy.to_cube(seg)
    n_t = seg["n_t"]
    dt = seg["dt_us"] * 1e-6
    mode = str(fx["vel_mode"])
    if mode == "const":
        prov = V.ConstantVelocity(float(str(fx["velocity_param"])))
    elif mode == "segy":
        vs = _read_bytes(fx["vel_file"])
        vc, vil, vxl = segy.to_cube(vs)
        prov = V.SegyVelocity(vc, vil, vxl, vs["dt_us"] * 1e-6)
    else:
        prov = V.TableVelocity(fx["vel_table"])
    vel = V.velocity_cube(prov, ill, xll, n_t, dt)
    c = fx["cfg"]
    num = dict(inline_step=float(c["inline_step"]), crossline_step=float(c["crossline_step"]),
               inline_padding=int(c.get("inline_padding", 0)),
               crossline_padding=int(c.get("crossline_padding", 0)),
               angle_aperture=float(c.get("angle_aperture", 30.0)),
               amp_correction=c.get("amp_correction", "true").lower() in ("true", "1", "yes", "on"))
    vmax = V.probe_max_velocity(prov, ill, xll, n_t, dt)
    return cube, vel, ill, xll, dt, num, vmax, seg


def output(fx):
    """The reference's output as [n_out_il][n_out_xl][n_t] in file order."""
    il, xl, tr = fx["out_inlines"], fx["out_crosslines"], fx["out_traces"]
    n_il = np.unique(il).size
    n_xl = np.unique(xl).size
    return tr.reshape(n_il, n_xl, -1), np.unique(il), np.unique(xl)
