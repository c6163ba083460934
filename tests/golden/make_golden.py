"""Writes the golden fixtures of tests/golden/ by running the reference CLI.

The reference `poststackmig` is compiled from its own sources by
`make -C oracle ref` into oracle/_ref/ (this script runs only where
/root/reference exists).  For each case it writes tiny synthetic SEG-Y inputs
with poststackmig3d_amd.segy, runs the reference binary on a config file, and
stores inputs, config and the reference's output file as one compressed .npz
(data only).  Re-run after changing a case:

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from poststackmig3d_amd import segy, synth  # noqa: E402
import pyoracle as O  # noqa: E402


def cube_case(name, cube, vel, cfg, il_labels=None, xl_labels=None, fmt=5, vel_mode="const",
              order=None, drop=None, vel_table=None, vel_dt_us=None, vel_cube=None):
    """vel_mode: 'const' (velocity=<number>), 'segy' (velocity SEG-Y on the
    data grid, or vel_cube with vel_dt_us), 'table' (text table).
    order: seed of a permutation of the trace order in the file; drop: trace
    numbers left out of the file."""
    n_il, n_xl, n_t = cube.shape
    ill = np.arange(1, n_il + 1) if il_labels is None else np.asarray(il_labels)
    xll = np.arange(1, n_xl + 1) if xl_labels is None else np.asarray(xl_labels)
    with tempfile.TemporaryDirectory() as td:
        inp = os.path.join(td, "in.sgy")
        out = os.path.join(td, "out.sgy")
        il = np.repeat(ill, n_xl)
        xl = np.tile(xll, n_il)
        traces = cube.reshape(-1, n_t)
        keep = np.ones(il.size, bool)
        if drop is not None:
            keep[np.asarray(drop)] = False
        il, xl, traces = il[keep], xl[keep], traces[keep]
        perm = None if order is None else np.random.default_rng(order).permutation(il.size)
        segy.write(inp, traces, il, xl, fmt=fmt, order=perm)
        lines = [f"input_data={inp}", f"output_data={out}"]
        store = {}
        if vel_mode == "const":
            lines.append(f"velocity={cfg['velocity']}")
        elif vel_mode == "segy":
            vp = os.path.join(td, "vel.sgy")
            vc = vel if vel_cube is None else vel_cube
            segy.write_cube(vp, vc, ill if vel_cube is None else None,
                            xll if vel_cube is None else None,
                            dt_us=vel_dt_us or 2000)
            lines.append(f"velocity={vp}")
            store["vel_file"] = np.fromfile(vp, np.uint8)
        else:
            vp = os.path.join(td, "vel.txt")
            with open(vp, "w") as f:
                f.write("# INLINE CROSSLINE TIME VEL\n")
                for row in vel_table:
                    f.write(" ".join(str(v) for v in row) + "\n")
            lines.append(f"velocity={vp}")
            store["vel_table"] = np.asarray(vel_table, np.float64)
        for k in ("inline_step", "crossline_step", "inline_padding", "crossline_padding",
                  "angle_aperture", "amp_correction", "n_threads"):
            if k in cfg:
                lines.append(f"{k}={cfg[k]}")
        cfgp = os.path.join(td, "config.txt")
        with open(cfgp, "w") as f:
            f.write("\n".join(lines) + "\n")
        r = subprocess.run([O.REF_CLI, cfgp], capture_output=True, text=True, check=True)
        res = segy.read(out)
        np.savez_compressed(
            os.path.join(HERE, f"{name}.npz"),
            input_file=np.fromfile(inp, np.uint8),
            output_file=np.fromfile(out, np.uint8),
            out_traces=res["traces"], out_inlines=res["inlines"], out_crosslines=res["crosslines"],
            config="\n".join(l for l in lines if not l.startswith(("input_data", "output_data",
                                                                   "velocity"))),
            velocity_param=str(cfg.get("velocity", vel_mode)), vel_mode=vel_mode,
            stdout=r.stdout, **store)
        print(name, res["traces"].shape, "ok")


def main():
    if not os.path.exists(O.REF_CLI):
        sys.exit("oracle/_ref/poststackmig is not built (make -C oracle ref)")
    # 1. plain method, constant velocity, no padding
    d, v = synth.point_diffractor(15, 13, 64, v=2000.0, noise=0.1, seed=21, t_apex=0.06)
    cube_case("g1_const_v", d, v, dict(velocity="2000.0", inline_step=25, crossline_step=25,
                                       angle_aperture=30, amp_correction="true"))
    # 2. padding both axes, laterally varying Vrms from a SEG-Y cube on the data grid
    n = (12, 11, 80)
    d, _ = synth.production(*n, n_diffractors=3, seed=22)
    v = synth.production_velocity(*n)
    cube_case("g2_pad_segy_vel", d, v, dict(inline_step=25, crossline_step=20,
                                           inline_padding=2, crossline_padding=3,
                                           angle_aperture=45, amp_correction="true"),
              vel_mode="segy")
    # 3. amplitude correction off, 60 degrees, velocity from a text table
    d, v = synth.point_diffractor(9, 10, 48, v=1800.0, noise=0.2, seed=23, t_apex=0.04)
    table = [(1, 1, 0.0, 1700.0), (1, 10, 0.05, 1900.0), (9, 1, 0.09, 2100.0),
             (9, 10, 0.02, 1800.0), (5, 5, 0.05, 2000.0)]
    cube_case("g3_table_vel_noamp", d, v, dict(inline_step=20, crossline_step=25,
                                              inline_padding=1, crossline_padding=1,
                                              angle_aperture=60, amp_correction="false"),
              vel_mode="table", vel_table=table)
    # 4. IBM float input, padded
    d, v = synth.point_diffractor(10, 9, 40, v=2200.0, noise=0.3, seed=24, t_apex=0.03)
    cube_case("g4_ibm_float", d, v, dict(velocity="2200", inline_step=25, crossline_step=25,
                                         crossline_padding=2, angle_aperture=40,
                                         amp_correction="on"), fmt=1)
    # 5. shuffled trace order, missing traces, non-unit label steps
    d, v = synth.point_diffractor(8, 9, 40, v=2000.0, noise=0.3, seed=25, t_apex=0.03)
    cube_case("g5_shuffled_missing", d, v, dict(velocity="2000", inline_step=25,
                                                crossline_step=25, inline_padding=1,
                                                crossline_padding=1, angle_aperture=35,
                                                amp_correction="yes"),
              il_labels=np.arange(100, 108) * 2, xl_labels=np.arange(10, 19) * 3,
              order=5, drop=[3, 17, 40])
    # 6. velocity SEG-Y on a coarser time grid (time and 2x2 lateral interpolation path)
    d, _ = synth.point_diffractor(7, 6, 40, v=2000.0, noise=0.2, seed=26, t_apex=0.03)
    vc = synth.production_velocity(7, 6, 20, dt=0.004)
    cube_case("g6_segy_vel_resampled", d, None, dict(inline_step=25, crossline_step=25,
                                                     angle_aperture=30,
                                                     amp_correction="true"),
              vel_mode="segy", vel_cube=vc, vel_dt_us=4000)


if __name__ == "__main__":
    main()
