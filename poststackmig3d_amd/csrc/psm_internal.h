// psm_internal.h -- shared declarations between the C ABI (psm_capi.cpp) and
// the gfx950 kernels (psm_kernels.hip).  Not installed; callers use psm.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "psm.h"

namespace psm {

// Kernel modes.
enum Mode : int {
  kForward = 0,  // migration: out = sum w lerp(x, q)
  kCount = 1,    // pair counts of the forward sum
  kAdjoint = 2,  // demigration: scatter the image into the data cube
  kTangent = 3,  // d(out)/d(alpha) for a velocity scale alpha, mask fixed
};

// One output inline of work: the centre row (in data-row index space), the
// data-row window [b0, b1) it sums over, and the velocity row it reads.
struct Task {
  int c, b0, b1, vrow;
};

// Arguments of the forward (migration) kernel.  Crossline coordinates are
// "buffer" coordinates as in migration_kernel.cpp: the data crosslines occupy
// [xd_off, xd_off + n_xd), the velocity crosslines [xv_off, xv_off + n_xv),
// and K is the buffer width n_buf_xl that clamps the crossline window.
struct MigArgs {
  const void* xp;          // traces in pair layout [n_rows + 1][n_xd + 1][n_t] of
                           // (x[i], x[i+1]-x[i]) in the `storage` format; sample n_t-1,
                           // the last trace of each row and the last row are zero
  const float* scale;      // [n_rows + 1][n_xd + 1] per-trace scale (F16 / BF16 storage)
  int n_rows;              // data rows (inlines) in xp
  int storage;             // PSM_STORE_*
  const float* vel;        // velocity rows [n_vrows][n_xv][n_t]
  float* out;              // [n_out_il][n_out_xl][n_t]
  const float* t0;         // [n_t]
  const Task* tasks;       // [n_out_il]
  const int* row_pos;      // generic-row mode: position of each data row; nullptr = contiguous
  unsigned long long* counters;  // count mode: {p_ap, p_g}
  float* out_g;            // count mode: per-sample P_g image (nullable; `out` holds P_ap)
  int n_t, n_xd, xd_off, n_xv, xv_off, K;
  int n_out_xl, cur_xl_start, n_out_il;
  int cpos;                // generic-row mode: position of the current inline
  float inv_dt, dx, dy, inv_dy, tan2;
  int amp;
  int d_lo, d_hi;          // this pass: row offsets |db| in [d_lo, d_hi) (generic rows: row indices)
  int accumulate;          // 0: out = sum; 1: out += sum (later passes)
  float alpha;             // velocity scale of travel time and weight (1 = reference)
  float alpha_mask;        // velocity scale of the aperture mask and window (1 = reference)
  const float* image;      // adjoint: image [n_out_il][n_out_xl][n_t] to scatter
  float* adj_out;          // adjoint: data [n_rows][n_xd][n_t], accumulated (zeroed by the caller)
  float* adj_rows;         // staged adjoint: row-pair windows [n_out_il][adj_nd][n_xd][n_t]
                           // at row offset d - d_lo, accumulated (zeroed by the caller) and
                           // added into both rows of each pair afterwards; nullptr = the
                           // windows are added into both data rows at once
  int adj_nd;              // row offsets adj_rows holds per output inline
};

// The batch work queue is split by XCD (workgroups are dispatched to the 8
// XCDs round robin by workgroup index): XCD x works through its own queue
// first, then helps the others, so the items each XCD runs at a time are
// neighbours that share aperture rows in its L2.
constexpr int kBatchXcds = 8;

// One segment of a batch's work queue: the crossline groups of one t0 tile of
// one volume (`count` items, spatial index io * n_jgroups + group).  Its items
// are split into kBatchXcds contiguous chunks; chunk x (spatial indices from
// xbase[x]) occupies positions [xfirst[x], xfirst[x + 1 segment]) of XCD x's
// queue.
struct BatchSeg {
  int item, tile, count, n_jgroups;
  int xbase[kBatchXcds];
  int xfirst[kBatchXcds];
};

// Launch-shape parameters (chosen by the autotuner or the caller).
// Row offsets per launch chosen by rows_per_pass = -1 for the gather form of
// FP32 traces: the FP32 gathers are bound by the cache hierarchy, and passes
// of 8 aperture rows keep the rows that concurrently running workgroups read
// within L2 (1.2x on the 401 x 401 x 1001 survey); half-precision gathers are
// VALU-bound and run the whole aperture in one launch, as the staged form
// does.
constexpr int kAutoRowsF32 = 8;

struct Tuning {
  int waves_per_block = 4;   // output crosslines per workgroup (1, 2, 4, 8, 16)
  int rows_per_pass = -1;    // aperture row offsets per launch; 0 = all in one launch,
                             // -1 = automatic (kAutoRowsF32 for FP32 traces, else 0)
  int staging = -1;          // forward migration: 1 = aperture windows staged in LDS
                             // (workgroups of 4 or 8 waves, contiguous rows), 0 = gathers,
                             // -1 = automatic (staged with kAutoStageWaves x kAutoStageSteps)
  int stage_steps = 4;       // crossline offsets per staged group (2 or 4)
};

// Automatic forward-kernel choice: windows staged in LDS by workgroups of 8
// output crosslines, 4 crossline offsets per staged group (both rows of a row
// pair summed into one FP32 window: 372 ms on the 401 x 401 x 1001 survey
// against 442 ms with 2 offsets), the whole aperture in one launch.  This
// beats the gather form, which is bound by the texture-address rate, for
// every storage form (FP16 traces: 387 against 468 ms).
constexpr int kAutoStageWaves = 8;
constexpr int kAutoStageSteps = 4;

// Bytes a row of the pair layout may span for the staged forms (the row is
// addressed through a buffer resource with a 32-bit size, and the staging
// loads mark lanes out of range with offsets just below 2^31); longer rows
// run the gather form.
constexpr size_t kLdsMaxRowBytes = (size_t)1 << 30;

// Launchers (psm_kernels.hip).  All enqueue on `stream` and return the HIP
// status of the launch.
// Bytes of the pair layout of n_rows x n_xd traces of n_t samples.
size_t pairs_bytes(int storage, int n_rows, int n_xd, int n_t);
// x: [n_rows][n_xd][n_t] raw traces -> xp: pair layout in `storage` format,
// scale: [n_rows + 1][n_xd + 1] per-trace scales (written for F16 / BF16).
hipError_t launch_pack(int storage, const float* x, void* xp, float* scale, int n_rows, int n_xd,
                       int n_t, hipStream_t stream);
// row_extent: largest row offset (generic rows: largest row index) any task
// can reach; the launcher splits [0, row_extent] into passes of
// tune.rows_per_pass row offsets.
hipError_t launch_migrate(const MigArgs& a, int mode, hipStream_t stream, const Tuning& tune,
                          int row_extent);
// Table velocity provider: vel[n_il][n_xl][n_t] from `n_pts` points
// (device memory), labels (device memory), minimum point velocity `vmin`.
hipError_t launch_table_velocity(const psm_velocity_point* pts, int n_pts, float vmin,
                                 const int32_t* il, const int32_t* xl, int n_il, int n_xl,
                                 int n_t, double dt, float* vel, hipStream_t stream);

hipError_t launch_fill_zero(float* p, size_t n, hipStream_t stream);
// Staged adjoint: true if launch_migrate runs the adjoint in the staged form
// for these arguments (contiguous rows, staging not switched off, rows within
// kLdsMaxRowBytes).
bool adjoint_staged(const MigArgs& a, const Tuning& tune);
// One launch of the staged adjoint over the row offsets [a.d_lo, a.d_hi).
hipError_t launch_adjoint_staged(const MigArgs& a, hipStream_t stream);
// data[r][k][i] += sum over e in [start[r], start[r + 1]) of
// rows[plane[e]][k][i] (rows: planes of n_xd x n_t floats): the row-pair
// windows of one staged-adjoint pass added into both rows of each pair, the
// planes of data row r listed in output-inline order.  start and plane are
// device memory.
hipError_t launch_adjoint_rows(const float* rows, float* data, const int* start, const int* plane,
                               int n_rows, int n_xd, int n_t, hipStream_t stream);
// Persistent batch launch (forward migration of several volumes): `blocks`
// workgroups of wpb waves drain the kBatchXcds queues described by segs;
// counters[x] (zero at launch) counts the positions taken from XCD x's queue,
// totals[x] its length (both device memory).  staged: workgroups of the
// staged form with wpb (4 or 8) output crosslines and stage_g (2 or 4)
// offsets per group, else the gather form with wpb waves per workgroup.
hipError_t launch_batch(int storage, bool staged, int stage_g, const MigArgs* items,
                        const BatchSeg* segs,
                        int n_segs, unsigned* counters, const unsigned* totals, int wpb,
                        int blocks, hipStream_t stream);
// out[0] = sum x[i] y[i] in float64; partial holds dot_scratch_doubles() - 1 doubles.
size_t dot_scratch_doubles();
hipError_t launch_dot_f64(const float* x, const float* y, size_t n, double* partial, double* out,
                          hipStream_t stream);

}  // namespace psm
