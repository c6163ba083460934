"""The batched velocity scan at full size against the reference kernel: the
point-diffractor cube at 301 x 301 x 1001 migrated under 16 Vrms scales
alpha = 0.85 .. 1.15 plus 8 sub-volumes from 64 x 64 x 501 to 301 x 301 x 1001,
all in one persistent graph-captured launch whose shape the autotuner picks,
with FP32 and FP16 trace storage.  Each item is checked on a bounded sample
(three output inlines x 16 crosslines, all samples) against the reference
kernel driven by main.cpp's loop on that item's own input: the cube (or
sub-cube) with Vrms multiplied by alpha in float32, which is the product the
kernel forms.  Bars: FP32 relative L2 <= 1e-5 per item and <= 1e-4 per trace;
FP16 storage <= 1e-3 per item against the same FP32 reference."""
import numpy as np
import pytest

import pyoracle as O

from poststackmig3d_amd.batch import SCAN_GEOMETRY, Batch, velocity_scan

pytestmark = pytest.mark.gpu


def _sample_rows(n):
    return np.array([n // 4, n // 2, (3 * n) // 4], np.int32)


def _reference_samples(items, base, vbase):
    """Reference output of every item on its sample: (rows, xl0, m, ref)."""
    fn = O.ref_migrate_sample if O.have_ref() else O.migrate_sample
    g = SCAN_GEOMETRY
    n0 = base.shape[0]
    out = []
    for prm, d, v, _, alpha in items:
        a, t = prm.n_il, prm.n_t
        o = (n0 - a) // 2
        dn = np.ascontiguousarray(base[o:o + a, o:o + a, :t])
        vn = np.ascontiguousarray(vbase[o:o + a, o:o + a, :t]) * np.float32(alpha)
        rows, m = _sample_rows(a), 16
        xl0 = a // 2 - m // 2
        ref = fn(dn, vn, g["dt"], g["inline_step"], g["crossline_step"], 0, 0,
                 g["angle_aperture"], True, rows, xl0, m)
        out.append((rows, xl0, m, ref))
    return out


def test_velocity_scan_against_reference():
    import torch
    print("scan: building 24 items (16 alpha x 301^2 x 1001 + 8 sub-volumes)", flush=True)
    items, base, vbase = velocity_scan(torch.device("cuda:0"))
    print("scan: reference samples", flush=True)
    refs = _reference_samples(items, base, vbase)
    for storage, bar in (("f32", 1e-5), ("f16", 1e-3)):
        b = Batch(items, storage=storage)
        try:
            w, bpc, ms = b.autotune()
            print(f"scan {storage}: autotuned waves/block {w}, blocks/CU {bpc}, {ms:.1f} ms",
                  flush=True)
            for it in items:
                it[3].fill_(float("nan"))
            b.run()
            b.ctx.synchronize()
        finally:
            b.close()
        worst_item = worst_trace = 0.0
        for k, ((prm, _, _, out, alpha), (rows, xl0, m, ref)) in enumerate(zip(items, refs)):
            got = out[torch.from_numpy(rows).long().cuda()][:, xl0:xl0 + m].cpu().numpy()
            assert np.isfinite(got).all(), k
            err = O.rel_l2(got, ref)
            g = got.reshape(-1, got.shape[-1]).astype(np.float64)
            r = ref.reshape(-1, ref.shape[-1]).astype(np.float64)
            tr = float((np.linalg.norm(g - r, axis=1) / np.linalg.norm(r, axis=1)).max())
            worst_item, worst_trace = max(worst_item, err), max(worst_trace, tr)
            assert err <= bar, (storage, k, alpha, err)
            if storage == "f32":
                assert tr <= 1e-4, (k, tr)
        print(f"scan {storage}: worst item rel-L2 {worst_item:.3e}, worst trace {worst_trace:.3e}",
              flush=True)
