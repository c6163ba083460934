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

The JSON line carries
  roofline      the dominant kernel (the forward migration kernel) against
                the bound its counters show: VALU issue.  Its VALU
                wave-instructions per launch come from rocprofv3 --pmc runs
                of this script (child processes after the timed steps), its
                duration from HIP events on its launch stream; HBM traffic
                (FETCH_SIZE / WRITE_SIZE, copy-calibrated), the HBM and L2
                rates and the gather-model bytes ride along.  Without a
                profiler the committed counter file is used only if it was
                taken on this kernel source (sha stamp); else the fields are
                null.
  parity        relative L2 of the GPU image against the reference kernel
                (compiled from its own sources into oracle/_ref) on the CPU
                sample below, and the worst output trace
  cpu_baseline  that reference kernel on a bounded sample of the same
                workload on this host's cores (>= 4 output inlines per
                thread, median of 3 timed runs after a warm-up)
  precision     FP16 / BF16 / FP32 trace storage against the FP64 oracle on
                the 101 x 101 x 501 point-diffractor cube (whole image), and
                their throughput on the production survey

  python bench.py [--steps K] [--warmup W] [--workload production|point|scan|lsm|e2e|dropin]
  python bench.py --workload scan     # batched velocity scan, one graph launch per step
  python bench.py --workload lsm      # 10-iteration CG least-squares migration, 201x201x751
  python bench.py --workload e2e      # bin/poststackmig on SEG-Y files of the production survey
  python bench.py --workload dropin   # per-inline drop-in calls at production window sizes
  python bench.py --copy-calibration   # plain 1 GiB device copy, for PMC scaling
"""
from __future__ import annotations

import argparse
import csv
import glob
import hashlib
import json
import os
import shutil
import statistics
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, 8 TB/s
N_CU = 256
PEAK_CLOCK_HZ = 2.4e9
# VALU issue peak: 256 CUs x 4 SIMDs, one wave64 instruction per 4 cycles each
VALU_PEAK = N_CU * 4 * PEAK_CLOCK_HZ / 4.0
PMC_FILE = os.path.join(ROOT, "profiles", "migrate_kernel_production_401_pmc.json")
KERNEL_SRC = os.path.join(ROOT, "poststackmig3d_amd", "csrc", "psm_kernels.hip")
# FETCH_SIZE / WRITE_SIZE (KiB per dispatch) -> bytes: a plain 1 GiB device copy
# reads 524 305 KiB FETCH_SIZE and writes 1 048 576 KiB WRITE_SIZE on MI355X
# (bench.py --copy-calibration under rocprofv3 --pmc)
FETCH_SCALE = 2.0
WRITE_SCALE = 1.0
# one counter group per profiler run (block limits: 8 SQ, 4 TCC, 4 TCP, 2 TA,
# 2 TD, 2 GRBM; FETCH_SIZE takes 3 TCC, WRITE_SIZE 2)
COUNTER_GROUPS = [
    ["FETCH_SIZE"],
    ["WRITE_SIZE", "TCC_HIT_sum", "TCC_MISS_sum", "TCP_TCC_READ_REQ_sum", "TA_TA_BUSY_sum",
     "TD_TD_BUSY_sum"],
    ["SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_INSTS_VMEM_RD",
     "SQ_BUSY_CYCLES", "GRBM_GUI_ACTIVE"],
]

PRODUCTION = dict(shape=(401, 401, 1001),
                  geo=dict(dt=0.002, inline_step=25.0, crossline_step=25.0, inline_padding=30,
                           crossline_padding=30, angle_aperture=45.0),
                  desc=("production 401x401x1001 @2ms, 25m bins, Vrms(il,xl,t)=1500+1200t"
                        "+150sin(2pi il/400)cos(2pi xl/400), 64 diffractors + 3 dipping "
                        "reflectors (30 Hz Ricker), pad 30/30, 45 deg, amp on"))
POINT = dict(shape=(101, 101, 501),
             geo=dict(dt=0.002, inline_step=25.0, crossline_step=25.0, inline_padding=0,
                      crossline_padding=0, angle_aperture=30.0),
             desc=("point diffractor 101x101x501 @2ms, 25m bins, V=2000, 25 Hz Ricker + 10% "
                   "noise, no padding, 30 deg, amp on"))


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["production", "point", "scan", "lsm", "e2e", "dropin"],
                    default="production")
    ap.add_argument("--storage", choices=["f32", "f16", "bf16"], default="f32",
                    help="trace storage of the main line (compute stays FP32)")
    ap.add_argument("--kernel", choices=["auto", "staged", "gather"], default="auto",
                    help="forward kernel form: the library's default (auto: staged, 8 "
                         "crosslines per workgroup, 4 offsets per group), aperture windows "
                         "staged in LDS (8 crosslines per workgroup, --stage-steps offsets per "
                         "group, one launch), or gathered through the vector memory path")
    ap.add_argument("--stage-steps", type=int, choices=[2, 4], default=4,
                    help="crossline offsets per staged group with --kernel staged")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the FP16/BF16 storage variants of the production line")
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU leg (baseline timing, parity and precision checks)")
    ap.add_argument("--no-counters", action="store_true",
                    help="do not run the rocprofv3 counter children")
    ap.add_argument("--cpu-rows", type=int, default=0,
                    help="output inlines in the CPU sample (0: 4 per host thread)")
    ap.add_argument("--cpu-xl", type=int, default=32, help="output crosslines per sampled inline")
    ap.add_argument("--counters-out", default=None,
                    help="also write the dominant kernel's counters (raw sums and derived "
                         "figures, stamped with the kernel source sha) to this JSON file")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--counter-child", action="store_true", help=argparse.SUPPRESS)
    return ap.parse_args()


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def kernel_source_sha():
    with open(KERNEL_SRC, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


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


# ---------------------------------------------------------------- host
def host_threads():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def host_info():
    """CPU model, sockets, cores and SMT of this host (from /proc/cpuinfo)."""
    info = {"model": None, "sockets": None, "cores": None, "logical_cpus": os.cpu_count(),
            "smt": None, "usable_cpus": host_threads()}
    try:
        procs, model = [], None
        cur = {}
        with open("/proc/cpuinfo") as f:
            for line in f:
                if not line.strip():
                    if cur:
                        procs.append(cur)
                    cur = {}
                    continue
                k, _, v = line.partition(":")
                cur[k.strip()] = v.strip()
        if cur:
            procs.append(cur)
        model = procs[0].get("model name") if procs else None
        sockets = {p.get("physical id") for p in procs}
        cores = {(p.get("physical id"), p.get("core id")) for p in procs}
        info.update(model=model, sockets=len(sockets), cores=len(cores),
                    smt=(len(procs) // max(1, len(cores))) if procs else None)
    except OSError:
        pass
    return info


# ---------------------------------------------------------------- counters
def run_counter_children(args, storage, kernel_form):
    """rocprofv3 --pmc runs of one migration step of this workload (child
    processes of this script), one counter group each.  Returns the summed
    counters of the forward-migration dispatches, or None."""
    prof = shutil.which("rocprofv3")
    if not prof:
        return None
    out = {}
    for group in COUNTER_GROUPS:
        d = tempfile.mkdtemp(prefix="psm_pmc_")
        cmd = [prof, "--pmc", *group, "-d", d, "-o", "run", "--output-format", "csv", "--",
               sys.executable, "-u", os.path.abspath(__file__), "--counter-child",
               "--workload", args.workload, "--storage", storage, "--kernel", kernel_form,
               "--stage-steps", str(args.stage_steps)]
        log("counters: " + " ".join(group))
        try:
            r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                               timeout=240)
            files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
            if r.returncode != 0 and not files:
                log(f"counters: rocprofv3 exited {r.returncode}: {r.stdout[-400:]}")
                return None
            vals, launches = {}, set()
            for fn in files:
                with open(fn) as f:
                    for row in csv.DictReader(f):
                        name = row.get("Kernel_Name", "")
                        if "migrate" not in name or "demigrate" in name:
                            continue
                        if "migrate_kernel" not in name and "migrate_lds_kernel" not in name:
                            continue
                        launches.add(row.get("Dispatch_Id"))
                        c = row.get("Counter_Name")
                        vals[c] = vals.get(c, 0.0) + float(row.get("Counter_Value", 0.0))
            missing = [c for c in group if c not in vals]
            if missing:
                log(f"counters: {missing} not found in the profile")
                return None
            out.update(vals)
            out["launches"] = len(launches)
        except (subprocess.TimeoutExpired, OSError, ValueError) as e:
            log(f"counters: {type(e).__name__}: {e}")
            return None
        finally:
            shutil.rmtree(d, ignore_errors=True)
    return out


def derive_counters(c, kernel_s):
    """Roofline quantities from the summed counters of one step's forward
    dispatches and the step's kernel time."""
    cycles = c["GRBM_GUI_ACTIVE"] / 8.0          # summed over 8 XCDs
    fetch = c["FETCH_SIZE"] * 1024.0 * FETCH_SCALE
    write = c["WRITE_SIZE"] * 1024.0 * WRITE_SCALE
    return {
        "valu_insts": c["SQ_INSTS_VALU"],
        "salu_insts": c["SQ_INSTS_SALU"],
        "lds_insts": c["SQ_INSTS_LDS"],
        "vmem_rd_insts": c["SQ_INSTS_VMEM_RD"],
        "waves": c["SQ_WAVES"],
        "clock_ghz": round(cycles / kernel_s / 1e9, 3) if kernel_s > 0 else None,
        "valu_busy": round(c["SQ_INSTS_VALU"] * 4.0 / (N_CU * 4 * cycles), 4),
        "salu_busy": round(c["SQ_INSTS_SALU"] * 4.0 / (N_CU * 4 * cycles), 4),
        "ta_busy": round(c["TA_TA_BUSY_sum"] / N_CU / cycles, 4),
        "td_busy": round(c["TD_TD_BUSY_sum"] / N_CU / cycles, 4),
        "l2_hit": round(c["TCC_HIT_sum"] / max(1.0, c["TCC_HIT_sum"] + c["TCC_MISS_sum"]), 4),
        "fetch_bytes": fetch,
        "write_bytes": write,
        "traffic_bytes": fetch + write,
        "l2_read_bytes": c["TCP_TCC_READ_REQ_sum"] * 64.0,
        "launches": c.get("launches"),
    }


def committed_counters(storage, form):
    """The committed counter file's entry, if it was taken on this kernel
    source with this storage and kernel form."""
    try:
        with open(PMC_FILE) as f:
            pmc = json.load(f)
    except (OSError, ValueError):
        return None, "no committed counter file"
    if pmc.get("kernel_sha16") != kernel_source_sha():
        return None, "committed counters are stale (kernel source changed)"
    ent = pmc.get(storage)
    if not ent or ent.get("form") != form:
        return None, f"no committed counters for {storage}/{form}"
    return ent, f"{os.path.relpath(PMC_FILE, ROOT)} (kernel sha {pmc['kernel_sha16']})"


def roofline_entry(derived, source, kernel_s, p_g, pair_bytes, kernel_name):
    gather = pair_bytes * p_g
    r = {"bound": "valu", "achieved": None, "peak": VALU_PEAK,
         "unit": "VALU wave64 instructions/s", "frac": None, "traffic": None,
         "kernel": kernel_name, "kernel_ms": round(kernel_s * 1e3, 3),
         "peak_is": "256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per wave64 VALU instruction",
         "counters_source": source}
    if derived:
        achieved = derived["valu_insts"] / kernel_s
        r.update(achieved=achieved, frac=round(achieved / VALU_PEAK, 4),
                 traffic=derived["traffic_bytes"],
                 hbm_gbs=round(derived["traffic_bytes"] / kernel_s / 1e9, 1),
                 hbm_frac=round(derived["traffic_bytes"] / kernel_s / 1e9 / HBM_PEAK_GBS, 4),
                 l2_read_gbs=round(derived["l2_read_bytes"] / kernel_s / 1e9, 1),
                 l2_read_is="TCP_TCC_READ_REQ_sum x 64 B (vector L1 -> L2 read requests)",
                 counters={k: derived[k] for k in ("valu_busy", "salu_busy", "ta_busy", "td_busy",
                                                   "l2_hit", "clock_ghz", "valu_insts",
                                                   "salu_insts", "lds_insts", "vmem_rd_insts")})
    r["gather_model"] = {
        "bytes": gather, "gbs": round(gather / kernel_s / 1e9, 1),
        "over_hbm_peak": round(gather / kernel_s / 1e9 / HBM_PEAK_GBS, 4),
        "note": f"{pair_bytes:.0f} B per gathering pair (P_g) as if every gather read HBM; "
                "not a roofline: each staged window is read once per workgroup and served "
                "from LDS to many output traces"}
    return r


# ---------------------------------------------------------------- CPU leg
def cpu_leg(d_np, v_np, geo, ap_img, g_img, gpu_img, rows_n, m):
    """The reference kernel (oracle/_ref; the oracle port where that is not
    built) on a bounded sample of the workload: `rows_n` evenly spaced output
    inlines (two of them padded when the survey is padded) x output
    crosslines [xl0, xl0 + m) centred in the image x all samples, on every
    usable host thread, median of 3 timed runs after a warm-up.  Returns the
    cpu_baseline entry and the parity of the GPU image on that sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle as O

    n_out_il, n_out_xl, _ = ap_img.shape
    pad = geo["inline_padding"]
    threads = host_threads()
    rows_n = min(rows_n or 4 * threads, n_out_il)
    inner = np.linspace(pad, n_out_il - 1 - pad, max(rows_n - (2 if pad else 0), 1)).round()
    rows = sorted(set(int(r) for r in inner))
    if pad:
        rows = sorted(set([pad // 2] + rows + [n_out_il - 1 - pad // 2]))
    rows = np.array(rows[:rows_n], np.int32)
    m = min(m, n_out_xl)
    xl0 = (n_out_xl - m) // 2
    p_ap = float(ap_img[rows][:, xl0:xl0 + m].astype(np.float64).sum())
    p_g = float(g_img[rows][:, xl0:xl0 + m].astype(np.float64).sum())
    kind = "reference" if O.have_ref() else "port"
    fn = O.ref_migrate_sample if kind == "reference" else O.migrate_sample
    args = (d_np, v_np, geo["dt"], geo["inline_step"], geo["crossline_step"],
            geo["inline_padding"], geo["crossline_padding"], geo["angle_aperture"], True)
    fn(*args, rows[:threads], xl0, 2, nthreads=threads)   # warm-up: page-in, OpenMP pool
    times, ref = [], None
    for k in range(3):
        t = time.perf_counter()
        ref = fn(*args, rows, xl0, m, nthreads=threads)
        times.append(time.perf_counter() - t)
        log(f"cpu baseline run {k + 1}/3: {times[-1]:.2f} s")
    dt = statistics.median(times)
    got = gpu_img[rows][:, xl0:xl0 + m]
    g = got.reshape(-1, got.shape[-1]).astype(np.float64)
    r = ref.reshape(-1, ref.shape[-1]).astype(np.float64)
    nz = np.linalg.norm(r, axis=1) > 0
    worst = float((np.linalg.norm(g - r, axis=1)[nz] / np.linalg.norm(r, axis=1)[nz]).max())
    parity = {"rel_l2": O.rel_l2(got, ref), "worst_trace_rel_l2": worst,
              "zero_where_reference_zero": bool(not g[~nz].any()),
              "against": f"{kind} kernel ({'oracle/_ref, compiled from its own sources' if kind == 'reference' else 'oracle restatement'}) on the cpu_baseline sample",
              "bar": 1e-5}
    cpu = {"value": p_ap / dt, "unit": "pairs/s", "cores": threads, "kind": kind,
           "value_p_g": p_g / dt,
           "sample": f"{len(rows)} output inlines (rows {rows.tolist()}) x {m} output "
                     f"crosslines [{xl0},{xl0 + m}) x all samples = {p_ap:.4e} P_ap / "
                     f"{p_g:.4e} P_g pairs; median of 3 runs {[round(x, 2) for x in times]} s "
                     f"after a warm-up; one output inline per OpenMP task, dynamic schedule "
                     f"(main.cpp:255)",
           "host": host_info()}
    return cpu, parity


def precision_leg(ctx_factory):
    """FP32 / FP16 / BF16 trace storage against the FP64 oracle on the
    101 x 101 x 501 point-diffractor cube (whole image)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle as O

    import poststackmig3d_amd as P
    from poststackmig3d_amd import synth
    data, vel = synth.point_diffractor(101, 101, 501)
    f64 = O.migrate_volume(data, vel, 0.002, 25.0, 25.0, 0, 0, 30.0, True, precision="f64")
    peak = float(np.abs(f64).max())
    prm = P.SurveyParams.for_cube(data.shape, angle_aperture=30.0)
    res = {}
    ctx = ctx_factory()
    try:
        for storage in ("f32", "f16", "bf16"):
            ctx.set_storage(storage)
            got = P.migrate_volume(data, vel, prm, ctx=ctx).astype(np.float64)
            res[storage] = {"rel_l2_vs_fp64": O.rel_l2(got, f64),
                            "max_err_over_peak": float(np.abs(got - f64).max() / peak)}
    finally:
        ctx.close()
    if O.have_ref():
        ref = O.ref_migrate_volume(data, vel, 0.002, 25.0, 25.0, 0, 0, 30.0, True)
        res["reference_f32"] = {"rel_l2_vs_fp64": O.rel_l2(ref, f64),
                                "max_err_over_peak": float(np.abs(ref - f64).max() / peak)}
    res["workload"] = "point diffractor 101x101x501, V=2000, 10% noise, 30 deg, amp on (whole image)"
    return res


# ---------------------------------------------------------------- other workloads
def run_scan(args, torch, P, dev):
    from poststackmig3d_amd.batch import Batch, velocity_scan
    ctx = P.context(0)
    items, _, _ = velocity_scan(dev)
    p_ap = p_g = 0.0
    for prm, d, v, out, alpha in items:
        ctx.set_velocity_scale(alpha, alpha)
        a, g = P.count_pairs(v, prm, ctx=ctx)
        p_ap += a
        p_g += g
    ctx.set_velocity_scale(1.0, 1.0)
    log(f"scan: {len(items)} items, P_ap={p_ap:.6e} P_g={p_g:.6e}")
    res = {}
    for storage in ("f32", "f16"):
        b = Batch(items, storage=storage, ctx=ctx)
        w, bpc, best = b.autotune()
        staged_form, _, _, g = b.shape()
        log(f"scan {storage}: autotuned waves/block={w} blocks/CU={bpc} "
            f"{'staged, %d offsets per group' % g if staged_form else 'gather form'} ({best:.1f} ms)")
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
                        "blocks_per_cu": bpc, "form": "staged" if staged_form else "gather",
                        "offsets_per_group": g}
        b.close()
    f32 = res["f32"]
    line = {
        "metric": "migrated sample-aperture-trace pairs/sec (1 GPU)",
        "value": f32["value"], "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(f32["ms_per_step"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded)",
        "config": {"workload": "velocity scan: point diffractor 301x301x1001 @2ms, V=2000 x "
                               "16 alpha in [0.85,1.15] + 8 sub-volumes 64x64x501..301x301x1001, "
                               "30 deg, amp on; one persistent graph-captured launch per step",
                   "items": len(items), "p_ap": p_ap, "p_g": p_g, "parallelism": "single GPU"},
        "roofline": None, "cpu_baseline": None,
        "storage_variants": {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv)
                                 for kk, vv in r.items()} for k, r in res.items()},
        "f16_over_f32": res["f16"]["value"] / f32["value"],
    }
    print(json.dumps(line), flush=True)
    return 0


def run_lsm(args, torch, P, dev):
    """Least-squares migration: 201 x 201 x 751 @ 2 ms, V = 2200 m/s, 30 deg,
    amp on; reflectivity = 0.5 % spikes ~ N(0, 1), data = demigration of it
    plus 5 % Gaussian noise.  One step = one 10-iteration CGLS solve
    (21 operator applications: F^T once, then F and F^T per iteration,
    FP64 dot products), checkpointed every 5 iterations.  A solve's last
    checkpoint is written behind the next solve; ms_per_step is the time of
    all K solves plus the wait for the last checkpoint to reach the disk,
    divided by K."""
    import numpy as np
    from poststackmig3d_amd import synth
    from poststackmig3d_amd.lsm import cgls, wait_checkpoints
    ctx = P.context(0)
    n = (201, 201, 751)
    vel = torch.full(n, 2200.0, device=dev)
    prm = P.SurveyParams.for_cube(n, angle_aperture=30.0, amp_correction=True)
    m_true = torch.from_numpy(synth.spikes(n, density=0.005, seed=3)).to(dev)
    d = P.demigrate_volume(m_true, vel, prm, ctx=ctx)
    g = torch.Generator(device=dev).manual_seed(5)
    d = d + 0.05 * d.std() * torch.randn(d.shape, device=dev, generator=g)
    p_ap, p_g = P.count_pairs(vel, prm, ctx=ctx)
    x = torch.randn(n, device=dev, generator=g)
    y = torch.randn(n, device=dev, generator=g)
    mx = P.migrate_volume(x, vel, prm, ctx=ctx)
    mty = P.demigrate_volume(y, vel, prm, ctx=ctx)
    dot_err = abs(P.dot_f64(mx, y, ctx=ctx) - P.dot_f64(x, mty, ctx=ctx)) / (
        np.sqrt(P.dot_f64(mx, mx, ctx=ctx)) * np.sqrt(P.dot_f64(y, y, ctx=ctx)))
    # one forward and one adjoint application, timed alone
    ctx.enable_timing(True)
    fwd_ms, adj_ms = [], []
    for _ in range(3):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        P.migrate_volume(x, vel, prm, out=mx, ctx=ctx)
        e1.record()
        e1.synchronize()
        fwd_ms.append(e0.elapsed_time(e1))
        torch.cuda.synchronize()
        e0.record()
        P.demigrate_volume(y, vel, prm, out=mty, ctx=ctx)
        e1.record()
        e1.synchronize()
        adj_ms.append(e0.elapsed_time(e1))
    ctx.enable_timing(False)
    del x, y, mx, mty
    log(f"lsm: P_ap={p_ap:.6e} per application, dot test {dot_err:.2e}, forward "
        f"{min(fwd_ms):.1f} ms, adjoint {min(adj_ms):.1f} ms")
    iters = 10
    applications = 1 + 2 * iters
    with tempfile.TemporaryDirectory() as tmp:
        ck = os.path.join(tmp, "cg.pt")
        for _ in range(args.warmup):
            cgls(d, vel, prm, iterations=2, ctx=ctx)
        times, res = [], None
        torch.cuda.synchronize()
        t_all = time.perf_counter()
        for i in range(args.steps):
            t0 = time.perf_counter()
            res = cgls(d, vel, prm, iterations=iters, ctx=ctx, checkpoint=ck, checkpoint_every=5)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
            log(f"lsm step {i + 1}/{args.steps}: {times[-1] * 1e3:.1f} ms, "
                f"||r|| {res.residual_norms[0]:.4e} -> {res.residual_norms[-1]:.4e}")
        t_drain = time.perf_counter()
        wait_checkpoints(ck)
        drain_ms = (time.perf_counter() - t_drain) * 1e3
        t_all = time.perf_counter() - t_all
        with open(ck, "rb"):
            pass   # the last solve's checkpoint is on disk
    mig = P.migrate_volume(d, vel, prm, ctx=ctx)
    corr = lambda u: float((u * m_true).sum() / (u.norm() * m_true.norm()))  # noqa: E731
    t = t_all / args.steps
    line = {
        "metric": "migrated sample-aperture-trace pairs/sec (1 GPU)",
        "value": p_ap * applications / t, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(t * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (FP64 dot products)",
        "data": "synthetic (seeded)",
        "config": {"workload": "CGLS least-squares migration, 10 iterations, 201x201x751 @2ms, "
                               "V=2200, 30 deg, amp on, 0.5% N(0,1) spikes, data = adjoint + 5% "
                               "noise, checkpoint every 5 iterations",
                   "applications_per_step": applications, "p_ap_per_application": p_ap,
                   "p_g_per_application": p_g, "parallelism": "single GPU"},
        "dot_test": dot_err,
        "solve_ms_median": round(float(np.median(times)) * 1e3, 3),
        "final_checkpoint_wait_ms": round(drain_ms, 3),
        "forward_ms": round(min(fwd_ms), 3), "adjoint_ms": round(min(adj_ms), 3),
        "residual_norms": [round(v, 6) for v in res.residual_norms],
        "reflectivity_correlation": {"migration": round(corr(mig), 4),
                                     "cgls_10": round(corr(res.model), 4)},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


def run_e2e(args, torch, P, dev):
    """bin/poststackmig on SEG-Y files of the production survey: input cube
    and Vrms written to IEEE SEG-Y in a temporary directory, then the whole
    command (ingest, velocity cube, geometry, migration, output writing) is
    timed as a child process, `steps` times."""
    from poststackmig3d_amd import segy, synth
    from poststackmig3d_amd.host_api import CLI_PATH
    shape, geo = PRODUCTION["shape"], PRODUCTION["geo"]
    tmp = tempfile.mkdtemp(prefix="psm_e2e_", dir=os.environ.get("TMPDIR", "/tmp"))
    try:
        data, vel = synth.production_torch(dev, *shape)
        d_np, v_np = data.cpu().numpy(), vel.cpu().numpy()
        del data, vel
        torch.cuda.empty_cache()
        t = time.perf_counter()
        segy.write_cube(os.path.join(tmp, "in.sgy"), d_np)
        segy.write_cube(os.path.join(tmp, "vel.sgy"), v_np)
        log(f"e2e: wrote 2 x {os.path.getsize(os.path.join(tmp, 'in.sgy')) / 1e9:.2f} GB SEG-Y "
            f"in {time.perf_counter() - t:.1f} s")
        del d_np, v_np
        cfg = os.path.join(tmp, "config.txt")
        with open(cfg, "w") as f:
            f.write(f"input_data = {tmp}/in.sgy\noutput_data = {tmp}/out.sgy\n"
                    f"velocity = {tmp}/vel.sgy\ninline_step = 25\ncrossline_step = 25\n"
                    f"inline_padding = 30\ncrossline_padding = 30\nangle_aperture = 45\n"
                    f"amp_correction = true\n")
        walls, migs = [], []
        for i in range(args.steps):
            t = time.perf_counter()
            r = subprocess.run([CLI_PATH, cfg], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                               text=True, timeout=600)
            walls.append(time.perf_counter() - t)
            if r.returncode != 0:
                log(r.stdout[-2000:])
                return r.returncode
            mig = [ln for ln in r.stdout.splitlines() if ln.startswith("Migration:")]
            migs.append(float(mig[0].split()[1]) if mig else None)
            log(f"e2e run {i + 1}/{args.steps}: {walls[-1]:.2f} s wall, migration {migs[-1]} s")
        out_bytes = os.path.getsize(os.path.join(tmp, "out.sgy"))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    wall = statistics.median(walls)
    line = {
        "metric": "end-to-end poststackmig run time (1 GPU)", "value": wall, "unit": "s",
        "n_gpus": 1, "steps": args.steps, "warmup": 0, "ms_per_step": round(wall * 1e3, 1),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded) SEG-Y files",
        "config": {"workload": PRODUCTION["desc"] + "; bin/poststackmig <config>: SEG-Y ingest, "
                               "SEG-Y Vrms cube, geometry, migration (with host transfers), "
                               "SEG-Y output", "output_bytes": out_bytes,
                   "parallelism": "single GPU"},
        "migration_s_incl_transfers": migs, "wall_s": [round(w, 3) for w in walls],
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


def run_dropin(args, torch, P, dev):
    """Per-inline drop-in (psm_kirchhoff_inline through the Python mirror) at
    the production survey's window sizes: the buffer of the centre output
    inline (375 x 461 x 1001 samples, 692 MB) and its velocity slice in host
    memory, as main.cpp hands them over; each call uploads, migrates one
    output inline and downloads it."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from poststackmig3d_amd import synth
    shape, geo = PRODUCTION["shape"], PRODUCTION["geo"]
    data, vel = synth.production_torch(dev, *shape)
    d_np, v_np = data.cpu().numpy(), vel.cpu().numpy()
    del data, vel
    prm = P.SurveyParams.for_cube(shape, amp_correction=True, **geo)
    info = P.survey_plan(prm, v_np)
    io = info.n_out_il // 2
    c = int(info.out_to_in_il[io])
    b0, b1 = (int(x) for x in info.window[io])
    pad = geo["crossline_padding"]
    K = shape[1] + 2 * pad
    buf = np.zeros((b1 - b0, K, shape[2]), np.float32)
    buf[:, pad:pad + shape[1]] = d_np[b0:b1]
    vs = np.zeros((K, shape[2]), np.float32)
    vs[pad:pad + shape[1]] = v_np[c]
    il_pos = np.arange(b0, b1, dtype=np.int32)
    t0 = info.t0_times
    out = np.empty((K, shape[2]), np.float32)
    ctx = P.context(0)
    times = []
    for i in range(args.warmup + args.steps):
        t = time.perf_counter()
        P.kirchhoff_kernel_3d_with_padding(buf, *buf.shape, vs, out, il_pos, None, c - b0, 0,
                                           geo["dt"], t0, info.tan_theta_sq, geo["inline_step"],
                                           geo["crossline_step"], True, ctx=ctx)
        if i >= args.warmup:
            times.append(time.perf_counter() - t)
            log(f"dropin call {len(times)}/{args.steps}: {times[-1] * 1e3:.1f} ms")
    per = statistics.median(times)
    line = {
        "metric": "per-inline drop-in call time (1 GPU)", "value": per * 1e3, "unit": "ms",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(per * 1e3, 2),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded)",
        "config": {"workload": PRODUCTION["desc"] + f"; one psm_kirchhoff_inline call for output "
                               f"inline {io} (buffer {b1 - b0}x{K}x{shape[2]} = "
                               f"{buf.nbytes / 1e6:.0f} MB host memory, pageable)",
                   "calls_per_volume": info.n_out_il, "parallelism": "single GPU"},
        "volume_estimate_s": round(per * info.n_out_il, 2),
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


# ---------------------------------------------------------------- main line
def setup_main(args, torch, P, dev):
    """Workload data resident in HBM, survey parameters and context."""
    from poststackmig3d_amd import synth
    wl = PRODUCTION if args.workload == "production" else POINT
    shape, geo = wl["shape"], wl["geo"]
    stream = torch.cuda.Stream(device=dev)
    ctx = P.context(0)
    ctx.set_stream(stream.cuda_stream)
    if args.kernel == "staged":
        ctx.set_tuning(8, 0)
        ctx.set_staging(True, args.stage_steps)
    elif args.kernel == "gather":
        ctx.set_staging(False)
    with torch.cuda.stream(stream):
        if args.workload == "production":
            data, vel = synth.production_torch(dev, *shape)
        else:
            d_np, v_np = synth.point_diffractor()
            data, vel = torch.from_numpy(d_np).to(dev), torch.from_numpy(v_np).to(dev)
    torch.cuda.synchronize()
    prm = P.SurveyParams.for_cube(shape, amp_correction=True, **geo)
    info = P.survey_plan(prm, vel)
    prm.max_velocity = info.max_velocity           # estimateMaxVelocity once, as main.cpp
    ctx.set_storage(args.storage)
    return wl, shape, geo, stream, ctx, data, vel, prm, info


def counter_child(args, torch, P, dev):
    """One migration step of the workload, for rocprofv3 --pmc."""
    wl, shape, geo, stream, ctx, data, vel, prm, info = setup_main(args, torch, P, dev)
    out = torch.empty((info.n_out_il, info.n_out_xl, info.n_t), dtype=torch.float32, device=dev)
    P.migrate_volume(data, vel, prm, out=out, ctx=ctx, sync=True)
    torch.cuda.synchronize()
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
    from poststackmig3d_amd._lib import Context

    dev = torch.device("cuda:0")
    if args.counter_child:
        return counter_child(args, torch, P, dev)
    if args.workload == "scan":
        return run_scan(args, torch, P, dev)
    if args.workload == "lsm":
        return run_lsm(args, torch, P, dev)
    if args.workload == "e2e":
        return run_e2e(args, torch, P, dev)
    if args.workload == "dropin":
        return run_dropin(args, torch, P, dev)

    wl, shape, geo, stream, ctx, data, vel, prm, info = setup_main(args, torch, P, dev)
    log(f"workload {wl['desc']}; output {info.n_out_il}x{info.n_out_xl}x{info.n_t}, "
        f"ap={info.max_aperture_il}, vmax={info.max_velocity}")
    p_ap, p_g, ap_img, g_img = P.count_pairs(vel, prm, ctx=ctx, images=True)
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
    gpu_img = out.cpu().numpy()

    staged = args.kernel != "gather"
    form = "staged" if staged else "gather"
    kernel_name = "%s (forward, %s traces)" % ("migrate_lds_kernel" if staged else "migrate_kernel",
                                               args.storage.upper())
    pair_bytes = 8.0 if args.storage == "f32" else 4.0

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
        ctx.set_storage(args.storage)
    del out
    torch.cuda.empty_cache()

    # counters of the dominant kernel: live rocprofv3 children, else the
    # committed file when it matches this kernel source
    derived, source = None, None
    if not args.no_counters:
        raw = run_counter_children(args, args.storage, args.kernel)
        if raw:
            derived = derive_counters(raw, avg_kernel_s)
            source = ("live: rocprofv3 --pmc child runs of this script (one migration step each, "
                      f"kernel sha {kernel_source_sha()})")
            if args.counters_out:
                rec = {}
                if os.path.exists(args.counters_out):
                    with open(args.counters_out) as f:
                        rec = json.load(f)
                if rec.get("kernel_sha16") != kernel_source_sha():
                    rec = {}
                rec.update({"kernel_sha16": kernel_source_sha(), "workload": args.workload,
                            "fetch_scale": FETCH_SCALE, "write_scale": WRITE_SCALE,
                            "counter_groups": COUNTER_GROUPS})
                rec[args.storage] = dict(derived, form=form, kernel_ms=round(avg_kernel_s * 1e3, 3),
                                         raw=raw, p_ap=p_ap, p_g=p_g,
                                         valu_insts_per_pair=derived["valu_insts"] / p_ap)
                with open(args.counters_out, "w") as f:
                    json.dump(rec, f, indent=1)
    if derived is None:
        ent, source = committed_counters(args.storage, form)
        if ent:
            derived = ent
    roofline = roofline_entry(derived, source, avg_kernel_s, p_g, pair_bytes, kernel_name)

    cpu = parity = precision = None
    if not args.no_cpu:
        d_np = data.cpu().numpy()
        v_np = vel.cpu().numpy()
        cpu, parity = cpu_leg(d_np, v_np, geo, ap_img.cpu().numpy(), g_img.cpu().numpy(),
                              gpu_img, args.cpu_rows, args.cpu_xl)
        log(f"cpu baseline ({cpu['kind']}, {cpu['cores']} threads): {cpu['value']:.4e} pairs/s; "
            f"parity rel-L2 {parity['rel_l2']:.3e} (worst trace {parity['worst_trace_rel_l2']:.3e})")
        precision = precision_leg(lambda: Context(0))
        for k, v in variants.items():
            if k in precision:
                precision[k]["production_over_f32"] = v["over_f32"]
        log("precision vs FP64: " + ", ".join(
            f"{k} {v['rel_l2_vs_fp64']:.2e}" for k, v in precision.items() if isinstance(v, dict)))

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
        "config": {"workload": wl["desc"], "n_il": shape[0], "n_xl": shape[1], "n_t": shape[2],
                   "out": [info.n_out_il, info.n_out_xl, info.n_t], "p_ap": p_ap, "p_g": p_g,
                   "trace_storage": args.storage, "kernel_form": args.kernel,
                   "parallelism": "single GPU"},
        "value_p_g": p_g / (wall / args.steps),
        "kernel_ms": round(avg_kernel_s * 1e3, 3),
        "kernel_ms_median": round(statistics.median(kernel_ms), 3),
        "step_ms_median": round(statistics.median(step_ms), 3),
        "pack_ms": round(sum(pack_ms) / len(pack_ms), 3),
        "roofline": roofline,
        "parity": parity,
        "cpu_baseline": cpu,
        "precision": precision,
        "storage_variants": variants or None,
        "gpu_over_cpu": (value / cpu["value"]) if cpu and cpu["value"] > 0 else None,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
