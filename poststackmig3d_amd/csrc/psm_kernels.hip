// psm_kernels.hip -- gfx950 kernels of the Kirchhoff post-stack 3D time
// migration (diffraction summation).
//
// Forward operator, per output sample (image point j at time t0) of one
// output inline, as MigrationKernel::kirchhoffKernel3DWithPadding
// (migration_kernel.cpp:17-383) defines it:
//
//   out = sum over aperture traces (b, k) with r^2 <= R^2 and i < n_t-1 of
//         w * lerp(x_bk, q),   q = tau/dt, i = trunc(q), f = q - i,
//         tau = sqrt(t0^2 + 4 r^2 / v^2),  R^2 = (v t0 / 2)^2 tan^2(theta),
//         w = 1/(tau v^2) when amp_correction, else 1,
//   r^2 = (db dx)^2 + (dk dy)^2, v = Vrms at the image point.
//
// Mapping to CDNA4:
//   * one wavefront = one output trace x 64 consecutive t0 samples (lane =
//     sample), so the aperture loop is wave-uniform and r^2 lives in scalar
//     registers; the 64 gathers of one aperture trace touch one contiguous
//     stretch of that trace (dtau/dt0 <= 1);
//   * the aperture is walked in (|db|, |dk|) order and each (|db|, |dk|)
//     shares tau, i, f and the weight between the up-to-four traces
//     (c +- db, kj +- dk) that have the same r^2 -- the travel time depends
//     only on r^2 and on the image point's velocity;
//   * traces are stored in a pair layout (x[i], x[i+1]-x[i]) so that one
//     8-byte load and one FMA give the linear interpolation of
//     migration_kernel.cpp:544-550, and the last sample holds (0,0) so the
//     bound i < n_t-1 (:298) becomes a clamp instead of a mask;
//   * workgroups are ordered latest t0 tile first (work grows as t0^2) and
//     striped so that each XCD takes a contiguous slab of output traces of
//     every tile (its own L2 sees neighbouring apertures).
//
// Arithmetic follows the reference's compiled order: r^2 = fma(dk dy, dk dy,
// dx^2) and tau^2 = fma(4/v^2, r^2, t0^2) are single-rounding FMAs (the
// -ffast-math build contracts them), tau is the correctly rounded square root,
// q = tau * (1/dt) with 1/dt rounded once.  The file is compiled with
// -ffp-contract=off so that these FMAs are exactly the ones written here.
#include <hip/hip_runtime.h>

#include "psm_internal.h"

namespace psm {

namespace {

constexpr int kWave = 64;
constexpr int kXcds = 8;

// Correctly rounded square root for positive normal x: v_sqrt_f32 is within
// one ulp, and the neighbour whose square brackets x is chosen by the sign of
// the fused residuals x - s_dn*s and x - s_up*s.
__device__ __forceinline__ float sqrt_rn(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const int si = __float_as_int(s);
  const float s_dn = __int_as_float(si - 1);
  const float s_up = __int_as_float(si + 1);
  const float r_dn = __builtin_fmaf(-s_dn, s, x);
  const float r_up = __builtin_fmaf(-s_up, s, x);
  float r = (r_dn <= 0.0f) ? s_dn : s;
  r = (r_up > 0.0f) ? s_up : r;
  return r;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, kWave));
  return v;
}

__device__ __forceinline__ unsigned wave_sum_u32(unsigned v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

__device__ __forceinline__ float uniform_f(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Per-lane image-point parameters (migration_kernel.cpp:115-144), evaluated
// for the crossline `src` whose velocity the lane uses.
struct LaneParams {
  float t0sq, a, R2, v2;
  int wlo, whi;
};

// One (|db|, |dk|) step of the aperture sum: the travel time, sample index,
// fraction and weight shared by the up-to-four traces with the same r^2, and
// their four pair-layout gathers.
struct Step {
  float f, w;
  bool m;
  float2 x0, x1, x2, x3;
};

// Sums the aperture of one (output trace, t0 tile) over the row offsets
// |db| in [d_lo, d_hi) of this pass.  WINDOW = some lane uses parameters of
// an earlier crossline whose crossline window can cut the aperture (sticky
// samples of migration_kernel.cpp:64-152), so each gather is also checked
// against that window.  COUNT = count pairs instead of summing.  All loop
// bounds and trace pointers are wave-uniform (scalar registers); absent
// partners read a trace of zeros so that the four gathers of a step are
// issued back to back, and the fast path loads step dk+1 before it consumes
// step dk (software pipelining: two steps of gathers in flight per wave).
template <bool GENERIC_ROWS, bool WINDOW, bool COUNT>
__device__ __forceinline__ void sum_aperture(const MigArgs& a, const Task& tk, int kj,
                                             const LaneParams& p, float R2max,
                                             float& acc, unsigned& cap, unsigned& cg) {
  const int n_t = a.n_t;
  const int nt1 = n_t - 1;
  const int xd_lo = a.xd_off, xd_hi = a.xd_off + a.n_xd;
  // pair layout rows are [n_xd + 1][n_t]: the last trace of every row is zero
  const size_t tr_bytes = (size_t)n_t * sizeof(float2);
  const size_t row_bytes = (size_t)(a.n_xd + 1) * tr_bytes;
  const char* const zero = reinterpret_cast<const char*>(a.zero_tr);
  const char* const base = reinterpret_cast<const char*>(a.xp);
  const char* const zero_row = reinterpret_cast<const char*>(a.zero_row);
  // largest crossline offset that can hold data
  const int dk_lim = max(kj - xd_lo, xd_hi - 1 - kj);

  // rowp / rowm: the rows at +db and -db (present when hp / hm)
  auto row_sweep = [&](const char* rowp, bool hp, const char* rowm, bool hm, float d2) {
    // last dk with r^2 <= R2max (r^2 is nondecreasing in dk)
    int kmax = (int)(__builtin_sqrtf(R2max - d2) * a.inv_dy) + 2;
    kmax = min(kmax, dk_lim);
    while (kmax >= 0) {
      const float dyv = (float)kmax * a.dy;
      if (__builtin_fmaf(dyv, dyv, d2) <= R2max) break;
      --kmax;
    }
    kmax = uniform_i(kmax);

    auto r2_of = [&](int dk) {
      const float dyv = (float)dk * a.dy;
      return uniform_f(__builtin_fmaf(dyv, dyv, d2));
    };

    if constexpr (!WINDOW && !COUNT) {
      // Fast path.  Each row is read through a buffer resource whose base is
      // the row; a gather's address is (scalar trace offset) + (lane sample
      // offset), so a step costs a few scalar adds and no 64-bit address
      // arithmetic.  Absent crosslines read the row's trailing zero trace,
      // absent rows the trailing zero row.
      const __amdgpu_buffer_rsrc_t rp = __builtin_amdgcn_make_buffer_rsrc(
          const_cast<char*>(rowp), 0, (int)row_bytes, 0x00020000);
      const __amdgpu_buffer_rsrc_t rm = __builtin_amdgcn_make_buffer_rsrc(
          const_cast<char*>(rowm), 0, (int)row_bytes, 0x00020000);
      const int tb = (int)tr_bytes;
      const int zero_off = a.n_xd * tb;                 // the row's zero trace
      const unsigned nxd = (unsigned)a.n_xd;
      auto prep = [&](int dk, Step& g) {
        const float dyv = (float)dk * a.dy;
        const float r2 = __builtin_fmaf(dyv, dyv, d2);
        const int xp_ = kj + dk - xd_lo, xm_ = kj - dk - xd_lo;   // data crossline indices
        const int op = ((unsigned)xp_ < nxd) ? xp_ * tb : zero_off;
        const int om = (dk > 0 && (unsigned)xm_ < nxd) ? xm_ * tb : zero_off;
        const float tau = sqrt_rn(__builtin_fmaf(p.a, r2, p.t0sq));
        const float q = tau * a.inv_dt;
        const int voff = min((int)q, nt1) * (int)sizeof(float2);
        g.x0 = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rp, voff, op, 0));
        g.x1 = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rp, voff, om, 0));
        g.x2 = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rm, voff, op, 0));
        g.x3 = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rm, voff, om, 0));
        g.f = __builtin_amdgcn_fractf(q);
        g.w = a.amp ? __builtin_amdgcn_rcpf(__builtin_fmaf(tau, p.v2, 1e-10f)) : 1.0f;
        g.m = r2 <= p.R2;
      };
      // a masked step adds 0 * w: the gathered sum is selected away before the
      // multiply, so a non-finite sample outside the aperture cannot reach acc
      auto consume = [&](const Step& g) {
        const float l0 = __builtin_fmaf(g.f, g.x0.y, g.x0.x);
        const float l1 = __builtin_fmaf(g.f, g.x1.y, g.x1.x);
        const float l2 = __builtin_fmaf(g.f, g.x2.y, g.x2.x);
        const float l3 = __builtin_fmaf(g.f, g.x3.y, g.x3.x);
        const float sum = g.m ? (l0 + l1) + (l2 + l3) : 0.0f;
        acc = __builtin_fmaf(sum, g.w, acc);
      };
      if (kmax < 0) return;
      // Two steps in flight: the gathers of step dk+1 are issued before step
      // dk is consumed (the scheduling barriers keep the compiler from
      // hoisting the consumption above them).  Steps past kmax are prefetched
      // and consumed unconditionally, so the load counter is the same on
      // every path; such a step either has r^2 > R2max (mask false in every
      // lane) or lies beyond the data on both sides (reads the zero trace).
      Step A, B;
      prep(0, A);
      for (int dk = 0; dk <= kmax; dk += 2) {
        prep(dk + 1, B);
        __builtin_amdgcn_sched_barrier(0);
        consume(A);
        prep(dk + 2, A);
        __builtin_amdgcn_sched_barrier(0);
        consume(B);
      }
    } else {
      for (int dk = 0; dk <= kmax; ++dk) {
        const float r2 = r2_of(dk);
        const int kp = kj + dk, km = kj - dk;
        const bool has_kp = kp >= xd_lo && kp < xd_hi;
        const bool has_km = dk > 0 && km >= xd_lo && km < xd_hi;
        if (!has_kp && !has_km) continue;
        const bool in_mask = r2 <= p.R2;
        const float tau = sqrt_rn(__builtin_fmaf(p.a, r2, p.t0sq));
        const float q = tau * a.inv_dt;
        const int i = (int)q;
        bool okp = true, okm = true;
        if (WINDOW) {
          okp = kp >= p.wlo && kp < p.whi;
          okm = km >= p.wlo && km < p.whi;
        }
        if constexpr (COUNT) {
          const unsigned nrow = (hp ? 1u : 0u) + (hm ? 1u : 0u);
          const unsigned n = nrow * ((has_kp && okp ? 1u : 0u) + (has_km && okm ? 1u : 0u));
          cap += in_mask ? n : 0u;
          cg += (in_mask && i < nt1) ? n : 0u;
        } else {
          const char* t0p = (hp && has_kp) ? rowp + (size_t)(kp - xd_lo) * tr_bytes : zero;
          const char* t0m = (hp && has_km) ? rowp + (size_t)(km - xd_lo) * tr_bytes : zero;
          const char* t1p = (hm && has_kp) ? rowm + (size_t)(kp - xd_lo) * tr_bytes : zero;
          const char* t1m = (hm && has_km) ? rowm + (size_t)(km - xd_lo) * tr_bytes : zero;
          const unsigned off = (unsigned)min(i, nt1) * (unsigned)sizeof(float2);
          const float2 x0 = *reinterpret_cast<const float2*>(t0p + off);
          const float2 x1 = *reinterpret_cast<const float2*>(t0m + off);
          const float2 x2 = *reinterpret_cast<const float2*>(t1p + off);
          const float2 x3 = *reinterpret_cast<const float2*>(t1m + off);
          const float f = __builtin_amdgcn_fractf(q);
          float w = a.amp ? __builtin_amdgcn_rcpf(__builtin_fmaf(tau, p.v2, 1e-10f)) : 1.0f;
          w = in_mask ? w : 0.0f;
          float l0 = __builtin_fmaf(f, x0.y, x0.x);
          float l1 = __builtin_fmaf(f, x1.y, x1.x);
          float l2 = __builtin_fmaf(f, x2.y, x2.x);
          float l3 = __builtin_fmaf(f, x3.y, x3.x);
          l0 = okp ? l0 : 0.0f;
          l2 = okp ? l2 : 0.0f;
          l1 = okm ? l1 : 0.0f;
          l3 = okm ? l3 : 0.0f;
          acc = __builtin_fmaf((l0 + l1) + (l2 + l3), w, acc);
        }
      }
    }
  };

  if constexpr (GENERIC_ROWS) {
    // generic rows: this pass covers the data rows [d_lo, d_hi)
    const int b_lo = max(tk.b0, a.d_lo), b_hi = min(tk.b1, a.d_hi);
    for (int b = b_lo; b < b_hi; ++b) {
      const float dxv = (float)(a.row_pos[b] - a.cpos) * a.dx;
      const float d2 = uniform_f(dxv * dxv);
      if (d2 > R2max) continue;
      row_sweep(base + (size_t)b * row_bytes, true, zero_row, false, d2);
    }
  } else {
    const int c = tk.c;
    int dmax = (int)(__builtin_sqrtf(R2max) / a.dx) + 2;
    dmax = min(dmax, max(tk.b1 - 1 - c, c - tk.b0));
    dmax = uniform_i(min(dmax, a.d_hi - 1));
    for (int d = max(0, a.d_lo); d <= dmax; ++d) {
      const float dxv = (float)d * a.dx;
      const float d2 = uniform_f(dxv * dxv);
      if (d2 > R2max) break;
      const int rp = c + d, rm = c - d;
      const bool hp = rp >= tk.b0 && rp < tk.b1;
      const bool hm = d > 0 && rm >= tk.b0 && rm < tk.b1;
      if (!hp && !hm) continue;
      row_sweep(hp ? base + (size_t)rp * row_bytes : zero_row, hp,
                hm ? base + (size_t)rm * row_bytes : zero_row, hm, d2);
    }
  }
}

template <int WPB, bool GENERIC_ROWS, bool COUNT>
__global__ __launch_bounds__(kWave * WPB) void migrate_kernel(MigArgs a, int n_tiles,
                                                              int n_jgroups, int per_tile_pad) {
  // ---- block -> (t0 tile, output inline, crossline group), heavy tiles first
  const int g = blockIdx.x;
  const int tile = n_tiles - 1 - g / per_tile_pad;
  const int rem = g % per_tile_pad;
  const int spatial = (rem % kXcds) * (per_tile_pad / kXcds) + rem / kXcds;
  if (spatial >= a.n_out_il * n_jgroups) return;
  const int io = spatial / n_jgroups;
  const int j = (spatial % n_jgroups) * WPB + uniform_i((int)(threadIdx.x / kWave));
  if (j >= a.n_out_xl) return;
  const int lane = threadIdx.x % kWave;
  const int s = tile * kWave + lane;
  const bool active = s < a.n_t;
  const int kj = a.cur_xl_start + j;
  float* orow = a.out + ((size_t)io * a.n_out_xl + j) * a.n_t;

  if (kj < 0 || kj >= a.K) {  // migration_kernel.cpp:79-85
    if (!COUNT && !a.accumulate && active) orow[s] = 0.0f;
    return;
  }
  const Task tk = a.tasks[io];
  const float* vrowp = a.vel + (size_t)tk.vrow * a.n_xv * a.n_t;

  // ---- parameter source: this crossline, or the last earlier crossline of
  // the call whose sample was valid (sticky caches, migration_kernel.cpp:64-152)
  const float t0 = active ? a.t0[s] : 0.0f;
  int src = -1;
  float v = 0.0f;
  if (active && t0 >= 1e-5f) {
    const int kmin = max(max(0, a.cur_xl_start), a.xv_off);
    for (int kk = min(kj, a.xv_off + a.n_xv - 1); kk >= kmin; --kk) {
      const float vv = vrowp[(size_t)(kk - a.xv_off) * a.n_t + s];
      if (vv >= 1e-5f) {
        src = kk;
        v = vv;
        break;
      }
    }
  }
  LaneParams p{0.0f, 0.0f, -1.0f, 1.0f, 0, 0};
  bool needs_window = false;
  if (src >= 0) {
    p.t0sq = t0 * t0;
    const float inv_v2 = 1.0f / (v * v + 1e-10f);
    p.a = 4.0f * inv_v2;
    const float vth = v * t0 * 0.5f;
    p.R2 = vth * vth * a.tan2;
    const int w = (int)(sqrt_rn(p.R2) * a.inv_dy) + 1;
    p.wlo = max(0, src - w);
    p.whi = min(a.K, src + w + 1);
    p.v2 = v * v;
    if (src != kj) {
      // crosslines this lane could reach around kj, clipped to the data
      const int lo_need = max(a.xd_off, kj - w);
      const int hi_need = min(a.xd_off + a.n_xd, kj + w + 1);
      needs_window = (p.wlo > lo_need) || (p.whi < hi_need);
    }
  }
  const float R2max = uniform_f(wave_max(p.R2));
  float acc = 0.0f;
  unsigned cap = 0, cg = 0;
  if (R2max >= 0.0f) {
    if (__any(needs_window))
      sum_aperture<GENERIC_ROWS, true, COUNT>(a, tk, kj, p, R2max, acc, cap, cg);
    else
      sum_aperture<GENERIC_ROWS, false, COUNT>(a, tk, kj, p, R2max, acc, cap, cg);
  }
  if constexpr (COUNT) {
    if (active && a.out) orow[s] = (float)cap;
    if (active && a.out_g) a.out_g[((size_t)io * a.n_out_xl + j) * a.n_t + s] = (float)cg;
    const unsigned tap = wave_sum_u32(cap), tg = wave_sum_u32(cg);
    if (lane == 0) {
      atomicAdd(&a.counters[0], (unsigned long long)tap);
      atomicAdd(&a.counters[1], (unsigned long long)tg);
    }
  } else if (active) {
    if (a.accumulate) {
      if (src >= 0) orow[s] += acc;
    } else {
      orow[s] = (src >= 0) ? acc : 0.0f;  // :148-151, :375
    }
  }
}

// Pair layout [n_rows + 1][n_xd + 1][n_t] of float2 (x[i], x[i+1] - x[i]):
// sample n_t-1 of every trace, the last trace of every row and the whole last
// row hold (0, 0).
__global__ void pack_pairs_kernel(const float* __restrict__ x, float2* __restrict__ xp,
                                  int n_rows, int n_xd, int n_t) {
  const size_t n = (size_t)(n_rows + 1) * (n_xd + 1) * n_t;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += stride) {
    const int i = (int)(idx % (size_t)n_t);
    const size_t tr = idx / (size_t)n_t;
    const int xl = (int)(tr % (size_t)(n_xd + 1));
    const int row = (int)(tr / (size_t)(n_xd + 1));
    float2 pr = make_float2(0.0f, 0.0f);
    if (row < n_rows && xl < n_xd && i < n_t - 1) {
      const size_t src = ((size_t)row * n_xd + xl) * n_t + i;
      const float lo = x[src];
      pr = make_float2(lo, x[src + 1] - lo);
    }
    xp[idx] = pr;
  }
}

}  // namespace

size_t pairs_elems(int n_rows, int n_xd, int n_t) {
  return (size_t)(n_rows + 1) * (n_xd + 1) * n_t;
}

hipError_t launch_pack_pairs(const float* x, float2* xp, int n_rows, int n_xd, int n_t,
                             hipStream_t stream) {
  const size_t n = pairs_elems(n_rows, n_xd, n_t);
  if (n == 0) return hipSuccess;
  const int threads = 256;
  size_t blocks = (n + threads - 1) / threads;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(pack_pairs_kernel, dim3((unsigned)blocks), dim3(threads), 0, stream, x, xp,
                     n_rows, n_xd, n_t);
  return hipGetLastError();
}

template <int WPB>
hipError_t launch_wpb(const MigArgs& a, bool count_only, hipStream_t stream) {
  const int n_tiles = (a.n_t + kWave - 1) / kWave;
  const int n_jgroups = (a.n_out_xl + WPB - 1) / WPB;
  const long long per_tile = (long long)a.n_out_il * n_jgroups;
  const long long per_tile_pad = (per_tile + kXcds - 1) / kXcds * kXcds;
  const long long blocks = per_tile_pad * n_tiles;
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  const dim3 grid((unsigned)blocks), block(kWave * WPB);
  const bool generic = a.row_pos != nullptr;
  const int ptp = (int)per_tile_pad;
  if (count_only) {
    if (generic)
      hipLaunchKernelGGL((migrate_kernel<WPB, true, true>), grid, block, 0, stream, a, n_tiles, n_jgroups, ptp);
    else
      hipLaunchKernelGGL((migrate_kernel<WPB, false, true>), grid, block, 0, stream, a, n_tiles, n_jgroups, ptp);
  } else {
    if (generic)
      hipLaunchKernelGGL((migrate_kernel<WPB, true, false>), grid, block, 0, stream, a, n_tiles, n_jgroups, ptp);
    else
      hipLaunchKernelGGL((migrate_kernel<WPB, false, false>), grid, block, 0, stream, a, n_tiles, n_jgroups, ptp);
  }
  return hipGetLastError();
}

hipError_t launch_migrate(const MigArgs& a0, bool count_only, hipStream_t stream,
                          const Tuning& tune, int row_extent) {
  if (a0.n_out_il <= 0 || a0.n_out_xl <= 0 || a0.n_t <= 0) return hipSuccess;
  MigArgs a = a0;
  // aperture-row passes: pass p sums the row offsets [p*R, (p+1)*R) (generic
  // rows: row indices); passes after the first accumulate into `out`
  const int R = (count_only || tune.rows_per_pass <= 0) ? (row_extent + 1) : tune.rows_per_pass;
  for (int lo = 0, pass = 0; lo <= row_extent; lo += R, ++pass) {
    a.d_lo = lo;
    a.d_hi = lo + R;
    a.accumulate = pass > 0 ? 1 : 0;
    hipError_t e;
    switch (tune.waves_per_block) {
      case 1: e = launch_wpb<1>(a, count_only, stream); break;
      case 2: e = launch_wpb<2>(a, count_only, stream); break;
      case 8: e = launch_wpb<8>(a, count_only, stream); break;
      case 16: e = launch_wpb<16>(a, count_only, stream); break;
      default: e = launch_wpb<4>(a, count_only, stream); break;
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace psm
