#!/usr/bin/env python3
"""Benchmark: Kirchhoff post-stack 3D time migration on one MI355X.

Workload (default): the production-size plain method -- 401 x 401 x 1001
samples at 2 ms, 25 m bins, Vrms(il,xl,t) = 1500 + 1200 t
+ 150 sin(2 pi il/400) cos(2 pi xl/400), 64 seeded point diffractors + 3
planar dipping reflectors (30 Hz Ricker), padding 30 traces per side, 45 deg
aperture, amplitude correction on, FP32.  Data are synthesised on the GPU and
stay resident in HBM; one step = one psm_migrate_volume over the whole cube
(trace packing + migration kernel), the unit of work being the
(output sample, aperture trace) pairs inside the aperture mask (P_ap),
counted exactly on the GPU.

Prints one JSON line with `roofline` (dominant kernel: the migration kernel,
algorithmic bytes = 8 B per gathering pair P_g, HIP-event duration on the
stream it runs on) and `cpu_baseline` (the reference's own kernel, compiled
from its sources into oracle/_ref, on a bounded sample of the same workload
on this host's cores).

  python bench.py [--steps K] [--warmup W] [--workload production|point|scan]
  python bench.py --workload scan     # batched velocity scan, one graph launch per step
  python bench.py --workload lsm      # 10-iteration CG least-squares migration, 201x201x751
  python bench.py --copy-calibration   # plain 1 GiB device copy, for PMC scaling
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, 8 TB/s
PMC_FILE = os.path.join(ROOT, "profiles", "migrate_kernel_production_401_pmc.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["production", "point", "scan", "lsm"],
                    default="production")
    ap.add_argument("--storage", choices=["f32", "f16", "bf16"], default="f32",
                    help="trace storage of the main line (compute stays FP32)")
    ap.add_argument("--kernel", choices=["auto", "staged", "gather"], default="auto",
                    help="forward kernel form: chosen by trace storage (auto), aperture "
                         "windows staged in LDS (8 crosslines per workgroup, 2 offsets per "
                         "group, one launch), or gathered through the vector memory path")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the FP16/BF16 storage variants of the production line")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-rows", type=int, default=16, help="output inlines in the CPU sample")
    ap.add_argument("--cpu-xl", type=int, default=192, help="output crosslines per sampled inline")
    ap.add_argument("--copy-calibration", action="store_true")
    return ap.parse_args()


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def copy_calibration():
    """One plain device-to-device copy of 1 GiB (for FETCH_SIZE/WRITE_SIZE
    scaling under rocprofv3 --pmc)."""
    import torch
    n = (1 << 30) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3}))


def cpu_threads():
    env = os.environ.get("OMP_NUM_THREADS")
    n = int(env) if env and env.isdigit() else (os.cpu_count() or 1)
    return max(1, min(n, os.cpu_count() or n))


def cpu_baseline(data_np, vel_np, geo, ap_img, rows_n, m):
    """The reference kernel (or, where oracle/_ref is not built, the oracle
    port) on a bounded sample: `rows_n` evenly spaced output inlines (two of
    them padded when the survey is padded), output crosslines [xl0, xl0+m)
    centred in the image.  Pairs of the sample come from the GPU count image."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle as O

    n_out_il, n_out_xl, _ = ap_img.shape
    pad = geo["inline_padding"]
    inner = np.linspace(pad, n_out_il - 1 - pad, max(rows_n - (2 if pad else 0), 1)).round()
    rows = list(inner.astype(int))
    if pad:
        rows = [pad // 2] + rows + [n_out_il - 1 - pad // 2]
    rows = np.array(sorted(set(rows))[:rows_n], np.int32)
    m = min(m, n_out_xl)
    xl0 = (n_out_xl - m) // 2
    pairs = float(ap_img[rows][:, xl0:xl0 + m].astype(np.float64).sum())
    threads = min(cpu_threads(), len(rows))
    kind = "reference" if O.have_ref() else "port"
    fn = O.ref_migrate_sample if kind == "reference" else O.migrate_sample
    args = (data_np, vel_np, geo["dt"], geo["inline_step"], geo["crossline_step"],
            geo["inline_padding"], geo["crossline_padding"], geo["angle_aperture"], True,
            rows, xl0, m)
    fn(*args[:9], rows[:1], xl0, 2, nthreads=1)  # warm-up (page-in, OpenMP pool)
    t = time.perf_counter()
    fn(*args, nthreads=threads)
    dt = time.perf_counter() - t
    return {"value": pairs / dt, "unit": "pairs/s", "cores": threads, "kind": kind,
            "sample": f"{len(rows)} output inlines (rows {rows.tolist()}) x {m} output "
                      f"crosslines [{xl0},{xl0 + m}) x all samples = {pairs:.4e} pairs in "
                      f"{dt:.2f} s, one inline per OpenMP thread (main.cpp:255 layout)",
            "seconds": dt}

SCAN_ALPHAS = 16
SCAN_SUBVOLS = 8


def scan_items(torch, dev, P, synth):
    """Velocity scan: the point-diffractor generator (V = 2000 m/s, 25 Hz
    Ricker at the cube centre, 10 % noise, 30 deg, amp on, no padding) at
    301 x 301 x 1001, migrated under 16 scales alpha = 0.85 .. 1.15 of its
    Vrms, plus 8 centred sub-volumes from 64 x 64 x 501 to 301 x 301 x 1001
    at alpha = 1."""
    import numpy as np
    d_np, v_np = synth.point_diffractor(301, 301, 1001)
    base = torch.from_numpy(d_np).to(dev)
    vbase = torch.from_numpy(v_np).to(dev)
    del d_np, v_np
    geo = dict(dt=0.002, inline_step=25.0, crossline_step=25.0, inline_padding=0,
               crossline_padding=0, angle_aperture=30.0, amp_correction=True)
    items = []
    for alpha in np.linspace(0.85, 1.15, SCAN_ALPHAS):
        prm = P.SurveyParams.for_cube(base.shape, **geo)
        items.append((prm, base, vbase, float(alpha)))
    n = np.linspace(64, 301, SCAN_SUBVOLS).round().astype(int)
    nt = np.linspace(501, 1001, SCAN_SUBVOLS).round().astype(int)
    for a, t in zip(n, nt):
        o = (301 - a) // 2
        d = base[o:o + a, o:o + a, :t].contiguous()
        v = vbase[o:o + a, o:o + a, :t].contiguous()
        items.append((P.SurveyParams.for_cube(d.shape, **geo), d, v, 1.0))
    return items


def run_scan(args, torch, P, synth, dev):
    from poststackmig3d_amd.batch import Batch
    ctx = P.context(0)
    items = scan_items(torch, dev, P, synth)
    full = []
    p_ap = 0.0
    p_g = 0.0
    for prm, d, v, alpha in items:
        info = P.survey_plan(prm, v)
        ctx.set_velocity_scale(alpha, alpha)
        a, g = P.count_pairs(v, prm, ctx=ctx)
        p_ap += a
        p_g += g
        out = torch.empty((info.n_out_il, info.n_out_xl, info.n_t), dtype=torch.float32,
                          device=dev)
        full.append((prm, d, v, out, alpha))
    ctx.set_velocity_scale(1.0, 1.0)
    log(f"scan: {len(full)} items, P_ap={p_ap:.6e} P_g={p_g:.6e}")
    res = {}
    for storage in ("f32", "f16"):
        b = Batch(full, storage=storage, ctx=ctx)
        w, bpc, best = b.autotune()
        log(f"scan {storage}: autotuned waves/block={w} blocks/CU={bpc} ({best:.1f} ms)")
        for _ in range(args.warmup):
            b.run()
        ms = []
        torch.cuda.synchronize()
        t_wall = time.perf_counter()
        for i in range(args.steps):
            ms.append(b.run(timed=True))
            log(f"scan {storage} step {i + 1}/{args.steps}: {ms[-1]:.2f} ms")
        torch.cuda.synchronize()
        wall = time.perf_counter() - t_wall
        res[storage] = {"value": p_ap / (wall / args.steps), "ms_per_step": wall * 1e3 / args.steps,
                        "graph_ms": sum(ms) / len(ms), "waves_per_block": w,
                        "blocks_per_cu": bpc}
        b.close()
    f32 = res["f32"]
    line = {
        "metric": "migrated sample-aperture-trace pairs/sec (1 GPU)",
        "value": f32["value"],
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(f32["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded)",
        "config": {"workload": "velocity scan: point diffractor 301x301x1001 @2ms, V=2000 x "
                               "16 alpha in [0.85,1.15] + 8 sub-volumes 64x64x501..301x301x1001, "
                               "30 deg, amp on; one persistent graph-captured launch per step",
                   "items": len(full), "p_ap": p_ap, "p_g": p_g, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(8.0 * p_g / (f32["graph_ms"] / 1e3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(8.0 * p_g / (f32["graph_ms"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "batch_kernel (graph incl. packs)"},
        "cpu_baseline": None,
        "storage_variants": {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv)
                                 for kk, vv in r.items()} for k, r in res.items()},
        "f16_over_f32": res["f16"]["value"] / f32["value"],
    }
    print(json.dumps(line), flush=True)
    return 0


def run_lsm(args, torch, P, synth, dev):
    """Least-squares migration: 201 x 201 x 751 @ 2 ms, V = 2200 m/s, 30 deg,
    amp on; reflectivity = 0.5 % spikes ~ N(0, 1), data = demigration of it
    plus 5 % Gaussian noise.  One step = one 10-iteration CGLS solve
    (21 operator applications: F^T once, then F and F^T per iteration,
    FP64 dot products), checkpointed every 5 iterations."""
    import tempfile
    import numpy as np
    from poststackmig3d_amd.lsm import cgls
    ctx = P.context(0)
    n = (201, 201, 751)
    vel = torch.full(n, 2200.0, device=dev)
    prm = P.SurveyParams.for_cube(n, angle_aperture=30.0, amp_correction=True)
    m_true = torch.from_numpy(synth.spikes(n, density=0.005, seed=3)).to(dev)
    d = P.demigrate_volume(m_true, vel, prm, ctx=ctx)
    g = torch.Generator(device=dev).manual_seed(5)
    d = d + 0.05 * d.std() * torch.randn(d.shape, device=dev, generator=g)
    p_ap, p_g = P.count_pairs(vel, prm, ctx=ctx)
    # dot-product test of the migration / demigration pair at this size
    x = torch.randn(n, device=dev, generator=g)
    y = torch.randn(n, device=dev, generator=g)
    mx = P.migrate_volume(x, vel, prm, ctx=ctx)
    mty = P.demigrate_volume(y, vel, prm, ctx=ctx)
    dot_err = abs(P.dot_f64(mx, y, ctx=ctx) - P.dot_f64(x, mty, ctx=ctx)) / (
        np.sqrt(P.dot_f64(mx, mx, ctx=ctx)) * np.sqrt(P.dot_f64(y, y, ctx=ctx)))
    del x, y, mx, mty
    log(f"lsm: P_ap={p_ap:.6e} per application, dot test {dot_err:.2e}")
    iters = 10
    applications = 1 + 2 * iters
    with tempfile.TemporaryDirectory() as tmp:
        ck = os.path.join(tmp, "cg.pt")
        for _ in range(args.warmup):
            cgls(d, vel, prm, iterations=2, ctx=ctx)
        times, res = [], None
        for i in range(args.steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            res = cgls(d, vel, prm, iterations=iters, ctx=ctx, checkpoint=ck, checkpoint_every=5)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
            log(f"lsm step {i + 1}/{args.steps}: {times[-1] * 1e3:.1f} ms, "
                f"||r|| {res.residual_norms[0]:.4e} -> {res.residual_norms[-1]:.4e}")
    mig = P.migrate_volume(d, vel, prm, ctx=ctx)
    corr = lambda u: float((u * m_true).sum() / (u.norm() * m_true.norm()))  # noqa: E731
    t = sum(times) / len(times)
    line = {
        "metric": "migrated sample-aperture-trace pairs/sec (1 GPU)",
        "value": p_ap * applications / t,
        "unit": "pairs/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (FP64 dot products)",
        "data": "synthetic (seeded)",
        "config": {"workload": "CGLS least-squares migration, 10 iterations, 201x201x751 @2ms, "
                               "V=2200, 30 deg, amp on, 0.5% N(0,1) spikes, data = adjoint + 5% "
                               "noise, checkpoint every 5 iterations",
                   "applications_per_step": applications, "p_ap_per_application": p_ap,
                   "p_g_per_application": p_g, "parallelism": "single GPU"},
        "dot_test": dot_err,
        "residual_norms": [round(v, 6) for v in res.residual_norms],
        "reflectivity_correlation": {"migration": round(corr(mig), 4),
                                     "cgls_10": round(corr(res.model), 4)},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return 0
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: measures one GPU only (the migration would shard by output "
              "inline ranges with no exchange; multi-GPU runs are not implemented).",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import poststackmig3d_amd as P
    from poststackmig3d_amd import synth

    dev = torch.device("cuda:0")
    if args.workload == "scan":
        return run_scan(args, torch, P, synth, dev)
    if args.workload == "lsm":
        return run_lsm(args, torch, P, synth, dev)
    stream = torch.cuda.Stream(device=dev)
    ctx = P.context(0)
    ctx.set_stream(stream.cuda_stream)
    if args.kernel == "staged":
        ctx.set_tuning(8, 0)
        ctx.set_staging(True, 2)
    elif args.kernel == "gather":
        ctx.set_staging(False)

    if args.workload == "production":
        shape = (401, 401, 1001)
        geo = dict(dt=0.002, inline_step=25.0, crossline_step=25.0, inline_padding=30,
                   crossline_padding=30, angle_aperture=45.0)
        desc = ("production 401x401x1001 @2ms, 25m bins, Vrms(il,xl,t)=1500+1200t"
                "+150sin(2pi il/400)cos(2pi xl/400), 64 diffractors + 3 dipping reflectors "
                "(30 Hz Ricker), pad 30/30, 45 deg, amp on")
        with torch.cuda.stream(stream):
            data, vel = synth.production_torch(dev, *shape)
    else:
        shape = (101, 101, 501)
        geo = dict(dt=0.002, inline_step=25.0, crossline_step=25.0, inline_padding=0,
                   crossline_padding=0, angle_aperture=30.0)
        desc = ("point diffractor 101x101x501 @2ms, 25m bins, V=2000, 25 Hz Ricker + 10% "
                "noise, no padding, 30 deg, amp on")
        d_np, v_np = synth.point_diffractor()
        data, vel = torch.from_numpy(d_np).to(dev), torch.from_numpy(v_np).to(dev)
    torch.cuda.synchronize()
    prm = P.SurveyParams.for_cube(shape, amp_correction=True, **geo)
    info = P.survey_plan(prm, vel)
    prm.max_velocity = info.max_velocity           # estimateMaxVelocity once, as main.cpp
    log(f"workload {desc}; output {info.n_out_il}x{info.n_out_xl}x{info.n_t}, "
        f"ap={info.max_aperture_il}, vmax={info.max_velocity}")

    ctx.set_storage(args.storage)
    p_ap, p_g, ap_img, _ = P.count_pairs(vel, prm, ctx=ctx, images=True)
    log(f"P_ap={p_ap:.6e} P_g={p_g:.6e}")
    out = torch.empty((info.n_out_il, info.n_out_xl, info.n_t), dtype=torch.float32, device=dev)

    for i in range(args.warmup):
        P.migrate_volume(data, vel, prm, out=out, ctx=ctx, sync=True)
        log(f"warmup {i + 1}/{args.warmup} done")
    ctx.enable_timing(True)
    kernel_ms, pack_ms, step_ms = [], [], []
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    for i in range(args.steps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        P.migrate_volume(data, vel, prm, out=out, ctx=ctx, sync=False)
        e1.record(stream)
        pk, mk = ctx.last_timing()
        e1.synchronize()
        kernel_ms.append(mk)
        pack_ms.append(pk)
        step_ms.append(e0.elapsed_time(e1))
        log(f"step {i + 1}/{args.steps}: {step_ms[-1]:.2f} ms (migrate kernel {mk:.2f} ms, "
            f"pack {pk:.3f} ms)")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    ctx.enable_timing(False)
    ms_per_step = wall * 1e3 / args.steps
    value = p_ap / (wall / args.steps)

    avg_kernel_s = sum(kernel_ms) / len(kernel_ms) / 1e3
    pair_bytes = 8.0 if args.storage == "f32" else 4.0   # one stored (x[i], x[i+1]-x[i]) pair
    alg_bytes = pair_bytes * p_g
    achieved = alg_bytes / avg_kernel_s / 1e9
    traffic = None
    counters = {}
    if os.path.exists(PMC_FILE):
        try:
            with open(PMC_FILE) as f:
                pmc = json.load(f)
            form = "staged" if (args.kernel == "staged" or
                                (args.kernel == "auto" and args.storage == "f32")) else "gather"
            ent = pmc.get(args.storage) if pmc.get("workload") == args.workload else None
            if ent and ent.get("form", "gather") == form:
                traffic = ent.get("traffic_bytes_per_step")
                counters = {k: ent[k] for k in ("valu_busy", "ta_busy", "td_busy", "l2_hit")
                            if k in ent}
        except (OSError, ValueError, KeyError):
            traffic = None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "%s (forward, FP%s traces)" % (
                    "migrate_lds_kernel" if (args.kernel == "staged" or
                                             (args.kernel == "auto" and args.storage == "f32"))
                    else "migrate_kernel", "32" if args.storage == "f32" else "16"),
                "kernel_ms": round(avg_kernel_s * 1e3, 3),
                "kernel_ms_is": "sum of the step's forward-kernel launches (one launch unless "
                                "aperture-row passes are set), HIP events on the launch stream",
                "algorithmic_bytes": alg_bytes,
                "rocprof_counters": counters or None,
                "traffic_gbs": round(traffic / avg_kernel_s / 1e9, 1) if traffic else None,
                "note": f"algorithmic bytes = {pair_bytes:.0f} B per gathering pair (P_g); the kernel reuses "
                        "each trace window across neighbouring output traces (from LDS in the "
                        "staged form, L1/L2 in the gather form), so HBM traffic is far below "
                        "this and frac > 1 is possible"}

    variants = {}
    if not args.no_variants and args.storage == "f32":
        for storage in ("f16", "bf16"):
            ctx.set_storage(storage)
            P.migrate_volume(data, vel, prm, out=out, ctx=ctx, sync=True)
            n = max(1, min(args.steps, 3))
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(n):
                P.migrate_volume(data, vel, prm, out=out, ctx=ctx, sync=False)
            torch.cuda.synchronize()
            w = (time.perf_counter() - t0) / n
            variants[storage] = {"value": p_ap / w, "ms_per_step": round(w * 1e3, 3),
                                 "over_f32": round((p_ap / w) / value, 3)}
            log(f"storage {storage}: {w * 1e3:.1f} ms/step, {variants[storage]['over_f32']:.2f}x f32")
        ctx.set_storage("f32")

    cpu = None
    if not args.no_cpu:
        d_np = data.cpu().numpy()
        v_np = vel.cpu().numpy()
        ap_np = ap_img.cpu().numpy()
        cpu = cpu_baseline(d_np, v_np, geo, ap_np, args.cpu_rows, args.cpu_xl)
        log(f"cpu baseline ({cpu['kind']}, {cpu['cores']} threads): {cpu['value']:.4e} pairs/s "
            f"in {cpu['seconds']:.1f} s")
        cpu.pop("seconds")

    line = {
        "metric": "migrated sample-aperture-trace pairs/sec (1 GPU)",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded, generated on the GPU)",
        "config": {"workload": desc, "n_il": shape[0], "n_xl": shape[1], "n_t": shape[2],
                   "out": [info.n_out_il, info.n_out_xl, info.n_t], "p_ap": p_ap, "p_g": p_g,
                   "trace_storage": args.storage, "kernel_form": args.kernel,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "storage_variants": variants or None,
        "gpu_over_cpu": (value / cpu["value"]) if cpu and cpu["value"] > 0 else None,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
