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
//     sample), so the aperture loop is wave-uniform and all trace addressing
//     is scalar; the 64 gathers of one aperture trace touch one contiguous
//     stretch of that trace (dtau/dt0 <= 1);
//   * the aperture is walked in (|db|, |dk|) order and each (|db|, |dk|)
//     shares tau, i, f and the weight between the up-to-four traces
//     (c +- db, kj +- dk) that have the same r^2 -- the travel time depends
//     only on r^2 and on the image point's velocity;
//   * traces are stored in a pair layout (x[i], x[i+1]-x[i]) so that one
//     gather and one FMA give the linear interpolation of
//     migration_kernel.cpp:544-550, and the last sample holds (0,0) so the
//     bound i < n_t-1 (:298) becomes a clamp instead of a mask.  The pair is
//     two float32 (8 bytes) or, for the reduced-precision storage forms, two
//     float16 / bfloat16 of the trace divided by its max |x| (4 bytes), with
//     the per-trace scale applied in float32;
//   * the gathers are buffer loads with a scalar trace offset, two steps in
//     flight per wave; they are bound by the texture-address issue rate,
//     which is why halving the pair width matters;
//   * workgroups are ordered latest t0 tile first (work grows as t0^2) and
//     striped so that each XCD takes a contiguous slab of output traces of
//     every tile; long apertures may be split into launches over a few
//     aperture rows each so the in-flight working set stays in L2.
//
// Arithmetic follows the reference's compiled order: r^2 = fma(dk dy, dk dy,
// dx^2) and tau^2 = fma(4/v^2, r^2, t0^2) are single-rounding FMAs (the
// -ffast-math build contracts them), tau is the correctly rounded square root,
// q = tau * (1/dt) with 1/dt rounded once.  The file is compiled with
// -ffp-contract=off so that these FMAs are exactly the ones written here.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "psm_internal.h"

namespace psm {

namespace {

constexpr int kWave = 64;
constexpr int kXcds = 8;
constexpr int kRsrcWord3 = 0x00020000;  // gfx9 buffer descriptor: 32-bit data format

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

// Read-only data indexed by wave-uniform values is read through the constant
// address space so that it goes to the scalar cache (s_load) instead of
// occupying the vector memory path.
using const_f32_ptr = const __attribute__((address_space(4))) float*;
__device__ __forceinline__ const_f32_ptr as_const(const float* p) { return (const_f32_ptr)p; }

// ---------------------------------------------------------------- storage
// One stored pair (x[i], x[i+1] - x[i]) and its linear interpolation in fp32.
template <int ST>
struct Store;

template <>
struct Store<PSM_STORE_F32> {
  using Raw = float2;
  static constexpr int kBytes = 8;
  static constexpr bool kScaled = false;
  __device__ static Raw load(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  }
  __device__ static Raw load_ptr(const char* p) { return *reinterpret_cast<const float2*>(p); }
  __device__ static float lerp(Raw x, float f) { return __builtin_fmaf(f, x.y, x.x); }
};

template <>
struct Store<PSM_STORE_F16> {
  using Raw = unsigned;
  static constexpr int kBytes = 4;
  static constexpr bool kScaled = true;
  __device__ static Raw load(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
  }
  __device__ static Raw load_ptr(const char* p) { return *reinterpret_cast<const unsigned*>(p); }
  __device__ static float lerp(Raw x, float f) {
    const _Float16 lo = __builtin_bit_cast(_Float16, (unsigned short)(x & 0xffffu));
    const _Float16 d = __builtin_bit_cast(_Float16, (unsigned short)(x >> 16));
    return __builtin_fmaf(f, (float)d, (float)lo);
  }
};

template <>
struct Store<PSM_STORE_BF16> {
  using Raw = unsigned;
  static constexpr int kBytes = 4;
  static constexpr bool kScaled = true;
  __device__ static Raw load(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
  }
  __device__ static Raw load_ptr(const char* p) { return *reinterpret_cast<const unsigned*>(p); }
  __device__ static float lerp(Raw x, float f) {
    return __builtin_fmaf(f, __uint_as_float(x & 0xffff0000u), __uint_as_float(x << 16));
  }
};

// Per-lane image-point parameters (migration_kernel.cpp:115-144), evaluated
// for the crossline `src` whose velocity the lane uses.
struct LaneParams {
  float t0sq, a, R2, v2;
  int wlo, whi;
  float mval;  // adjoint: the image sample this lane scatters
};

// Sums the aperture of one (output trace, t0 tile) over the row offsets
// |db| in [d_lo, d_hi) of this pass.  WINDOW = some lane uses parameters of
// an earlier crossline whose crossline window can cut the aperture (sticky
// samples of migration_kernel.cpp:64-152), so each gather is also checked
// against that window.  MODE: kForward sums, kCount counts pairs, kAdjoint
// scatters the lane's image sample into the data cube with the same
// (i, f, w, mask) as the forward sum, kTangent sums the derivative of each
// term with respect to the velocity scale.  All loop
// bounds and trace offsets are wave-uniform (scalar registers); absent
// crosslines read the row's zero trace and absent rows the zero row, so the
// four gathers of a step are issued back to back, and the fast path loads
// step dk+1 before it consumes step dk (two steps of gathers in flight).
template <int ST, bool GENERIC_ROWS, bool WINDOW, int MODE>
__device__ __forceinline__ void sum_aperture(const MigArgs& a, const Task& tk, int kj,
                                             const LaneParams& p, float R2max,
                                             float& acc, unsigned& cap, unsigned& cg) {
  using S = Store<ST>;
  constexpr bool COUNT = MODE == kCount;
  const int n_t = a.n_t;
  const int nt1 = n_t - 1;
  const int xd_lo = a.xd_off, xd_hi = a.xd_off + a.n_xd;
  // pair layout rows are [n_xd + 1][n_t]: the last trace of every row is zero
  const int tb = n_t * S::kBytes;
  const size_t row_bytes = (size_t)(a.n_xd + 1) * tb;
  const char* const base = reinterpret_cast<const char*>(a.xp);
  const char* const zero_row = base + (size_t)a.n_rows * row_bytes;
  const int row_tr = a.n_xd + 1;  // traces per row (scales)
  // largest crossline offset that can hold data
  const int dk_lim = max(kj - xd_lo, xd_hi - 1 - kj);

  // rowp / rowm: rows at +db and -db (present when hp / hm), with their
  // index for the scale table
  auto row_sweep = [&](const char* rowp, bool hp, int ip, const char* rowm, bool hm, int im,
                       float d2) {
    // last dk with r^2 <= R2max (r^2 is nondecreasing in dk)
    int kmax = (int)(__builtin_sqrtf(R2max - d2) * a.inv_dy) + 2;
    kmax = min(kmax, dk_lim);
    while (kmax >= 0) {
      const float dyv = (float)kmax * a.dy;
      if (__builtin_fmaf(dyv, dyv, d2) <= R2max) break;
      --kmax;
    }
    kmax = uniform_i(kmax);
    if (kmax < 0) return;
    const const_f32_ptr scp = as_const(a.scale + (size_t)ip * row_tr);
    const const_f32_ptr scm = as_const(a.scale + (size_t)im * row_tr);
    const unsigned nxd = (unsigned)a.n_xd;

    if constexpr (!WINDOW && MODE == kForward) {
      // Fast path.  Each present row is read through a buffer resource whose
      // base is the row; a gather's address is (scalar trace offset) + (lane
      // sample offset).  The crossline offsets dk in [0, kmax] are split into
      // at most three ranges by which of the columns kj + dk, kj - dk hold
      // data, and each range loads only the traces that exist: NR rows x NC
      // columns per step.  Absent traces would add exact zeros, so the sum is
      // the same as over all four partners.
      const char* row0 = hp ? rowp : rowm;
      const const_f32_ptr sc0 = hp ? scp : scm;
      const bool two_rows = hp && hm;
      const __amdgpu_buffer_rsrc_t rs0 =
          __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(row0), 0, (int)row_bytes, kRsrcWord3);
      const __amdgpu_buffer_rsrc_t rs1 =
          __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(rowm), 0, (int)row_bytes, kRsrcWord3);

      // Pipelined sweep of dk in [lo, hi] with NR rows and NC columns: the
      // columns are kj + sgn*dk (NC = 1) or kj + dk, kj - dk (NC = 2).  The
      // gathers of step dk+1 are issued before step dk is consumed; the
      // scheduling barriers keep the compiler from hoisting the consumption
      // above them.  A step prefetched past `hi` is never consumed; its
      // loads read the zero trace or lie inside the row.
      auto sweep = [&](auto nr_c, auto nc_c, int lo, int hi, int sgn) {
        constexpr int NR = decltype(nr_c)::value;
        constexpr int NC = decltype(nc_c)::value;
        struct G {
          float f, w;
          bool m;
          typename S::Raw x[NR * NC];
          float sc[NR * NC];
        };
        auto prep = [&](int dk, G& g) {
          const float dyv = (float)dk * a.dy;
          const float r2 = __builtin_fmaf(dyv, dyv, d2);
          int col[NC], off[NC];
          col[0] = kj + (NC == 2 ? dk : sgn * dk) - xd_lo;
          if constexpr (NC == 2) col[1] = kj - dk - xd_lo;
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            const bool ok = (unsigned)col[c] < nxd;
            col[c] = ok ? col[c] : a.n_xd;
            off[c] = col[c] * tb;
          }
          const float tau = sqrt_rn(__builtin_fmaf(p.a, r2, p.t0sq));
          const float q = tau * a.inv_dt;
          const int voff = min((int)q, nt1) * S::kBytes;
#pragma unroll
          for (int c = 0; c < NC; ++c) g.x[c] = S::load(rs0, voff, off[c]);
          if constexpr (NR == 2) {
#pragma unroll
            for (int c = 0; c < NC; ++c) g.x[NC + c] = S::load(rs1, voff, off[c]);
          }
          if constexpr (S::kScaled) {
#pragma unroll
            for (int c = 0; c < NC; ++c) g.sc[c] = sc0[col[c]];
            if constexpr (NR == 2) {
#pragma unroll
              for (int c = 0; c < NC; ++c) g.sc[NC + c] = scm[col[c]];
            }
          }
          g.f = __builtin_amdgcn_fractf(q);
          g.w = a.amp ? __builtin_amdgcn_rcpf(__builtin_fmaf(tau, p.v2, 1e-10f)) : 1.0f;
          g.m = r2 <= p.R2;
        };
        // a masked step adds 0 * w: the gathered sum is selected away before
        // the multiply, so a non-finite sample outside the aperture cannot
        // reach acc.  Slot order (row, column) is that of the four-partner
        // sum (+db,+dk), (+db,-dk), (-db,+dk), (-db,-dk) with absent slots
        // left out.
        auto consume = [&](const G& g) {
          float l[NR * NC];
#pragma unroll
          for (int k = 0; k < NR * NC; ++k) l[k] = S::lerp(g.x[k], g.f);
          float sum;
          if constexpr (S::kScaled) {
            sum = l[0] * g.sc[0];
#pragma unroll
            for (int k = 1; k < NR * NC; ++k) sum = __builtin_fmaf(l[k], g.sc[k], sum);
          } else if constexpr (NR * NC == 4) {
            sum = (l[0] + l[1]) + (l[2] + l[3]);
          } else if constexpr (NR * NC == 2) {
            sum = l[0] + l[1];
          } else {
            sum = l[0];
          }
          sum = g.m ? sum : 0.0f;
          acc = __builtin_fmaf(sum, g.w, acc);
        };
        if (lo > hi) return;
        G A, B;
        int dk = lo;
        prep(dk, A);
        for (; dk + 1 <= hi; dk += 2) {
          prep(dk + 1, B);
          __builtin_amdgcn_sched_barrier(0);
          consume(A);
          prep(dk + 2, A);
          __builtin_amdgcn_sched_barrier(0);
          consume(B);
        }
        if (dk <= hi) consume(A);
      };
      auto sweep_rows = [&](auto nc_c, int lo, int hi, int sgn) {
        if (two_rows)
          sweep(std::integral_constant<int, 2>{}, nc_c, lo, hi, sgn);
        else
          sweep(std::integral_constant<int, 1>{}, nc_c, lo, hi, sgn);
      };
      // dk ranges with data: +dk side [p_lo, p_hi], -dk side [m_lo, m_hi]
      const int p_lo = max(0, xd_lo - kj), p_hi = min(kmax, xd_hi - 1 - kj);
      const int m_lo = max(1, kj - (xd_hi - 1)), m_hi = min(kmax, kj - xd_lo);
      const int b_lo = max(p_lo, m_lo), b_hi = min(p_hi, m_hi);
      const std::integral_constant<int, 1> one{};
      const std::integral_constant<int, 2> two{};
      if (b_lo > b_hi) {  // no offset with both columns: the sides in dk order
        if (p_lo <= m_lo) {
          sweep_rows(one, p_lo, p_hi, 1);
          sweep_rows(one, m_lo, m_hi, -1);
        } else {
          sweep_rows(one, m_lo, m_hi, -1);
          sweep_rows(one, p_lo, p_hi, 1);
        }
      } else {
        if (p_lo < m_lo) sweep_rows(one, p_lo, b_lo - 1, 1);
        else sweep_rows(one, m_lo, b_lo - 1, -1);
        sweep_rows(two, b_lo, b_hi, 0);
        if (p_hi > m_hi) sweep_rows(one, b_hi + 1, p_hi, 1);
        else sweep_rows(one, b_hi + 1, m_hi, -1);
      }
    } else {
      for (int dk = 0; dk <= kmax; ++dk) {
        const float dyv = (float)dk * a.dy;
        const float r2 = __builtin_fmaf(dyv, dyv, d2);
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
        } else if constexpr (MODE == kAdjoint) {
          // d[b][k][i] += w (1 - f) m,  d[b][k][i+1] += w f m  (transpose of
          // the interpolation fma(f, x[i+1] - x[i], x[i]))
          if (in_mask && i < nt1) {
            const float f = __builtin_amdgcn_fractf(q);
            const float w = a.amp ? __builtin_amdgcn_rcpf(__builtin_fmaf(tau, p.v2, 1e-10f)) : 1.0f;
            const float wm = w * p.mval;
            const float c0 = wm - wm * f, c1 = wm * f;
            auto scatter = [&](bool row_ok, int row, bool k_ok, int k) {
              if (row_ok && k_ok) {
                float* d = a.adj_out + ((size_t)row * a.n_xd + (k - xd_lo)) * n_t + i;
                unsafeAtomicAdd(d, c0);
                unsafeAtomicAdd(d + 1, c1);
              }
            };
            scatter(hp, ip, has_kp && okp, kp);
            scatter(hp, ip, has_km && okm, km);
            scatter(hm, im, has_kp && okp, kp);
            scatter(hm, im, has_km && okm, km);
          }
        } else {
          const int xp_ = has_kp ? kp - xd_lo : a.n_xd;
          const int xm_ = has_km ? km - xd_lo : a.n_xd;
          const size_t off = (size_t)min(i, nt1) * S::kBytes;
          const float f = __builtin_amdgcn_fractf(q);
          const typename S::Raw x0 = S::load_ptr(rowp + (size_t)xp_ * tb + off);
          const typename S::Raw x1 = S::load_ptr(rowp + (size_t)xm_ * tb + off);
          const typename S::Raw x2 = S::load_ptr(rowm + (size_t)xp_ * tb + off);
          const typename S::Raw x3 = S::load_ptr(rowm + (size_t)xm_ * tb + off);
          float l0 = S::lerp(x0, f);
          float l1 = S::lerp(x1, f);
          float l2 = S::lerp(x2, f);
          float l3 = S::lerp(x3, f);
          if constexpr (S::kScaled) {
            l0 *= scp[xp_];
            l1 *= scp[xm_];
            l2 *= scm[xp_];
            l3 *= scm[xm_];
          }
          l0 = okp ? l0 : 0.0f;
          l2 = okp ? l2 : 0.0f;
          l1 = okm ? l1 : 0.0f;
          l3 = okm ? l3 : 0.0f;
          float sum = (l0 + l1) + (l2 + l3);
          sum = in_mask ? sum : 0.0f;
          const float w = a.amp ? __builtin_amdgcn_rcpf(__builtin_fmaf(tau, p.v2, 1e-10f)) : 1.0f;
          if constexpr (MODE == kTangent) {
            // d/d(alpha) of w * lerp(x, tau/dt) with v -> alpha v and the mask
            // held fixed: dtau = -(a r^2) / (alpha tau), dq = dtau / dt,
            // dw = -w (dtau / tau + 2 / alpha); d lerp / dq = x[i+1] - x[i]
            float dsum = 0.0f;
            if constexpr (!S::kScaled) {
              dsum = (okp ? x0.y + x2.y : 0.0f) + (okm ? x1.y + x3.y : 0.0f);
            }
            dsum = in_mask ? dsum : 0.0f;
            const float dtau = -(p.a * r2) / (a.alpha * tau);
            const float dq = dtau * a.inv_dt;
            const float dw = a.amp ? -w * (dtau / tau + 2.0f / a.alpha) : 0.0f;
            acc += dw * sum + w * dq * dsum;
          } else {
            acc = __builtin_fmaf(sum, w, acc);
          }
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
      row_sweep(base + (size_t)b * row_bytes, true, b, zero_row, false, a.n_rows, d2);
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
      row_sweep(hp ? base + (size_t)rp * row_bytes : zero_row, hp, hp ? rp : a.n_rows,
                hm ? base + (size_t)rm * row_bytes : zero_row, hm, hm ? rm : a.n_rows, d2);
    }
  }
}

// The work of one wave: output trace (io, j) of the output inline io, t0 tile
// `tile` (64 samples).
template <int ST, bool GENERIC_ROWS, int MODE>
__device__ __forceinline__ void migrate_trace_tile(const MigArgs& a, int tile, int io, int j) {
  constexpr bool COUNT = MODE == kCount;
  constexpr bool WRITES_OUT = MODE == kForward || MODE == kTangent;
  const int lane = threadIdx.x % kWave;
  const int s = tile * kWave + lane;
  const bool active = s < a.n_t;
  const int kj = a.cur_xl_start + j;
  float* orow = a.out + ((size_t)io * a.n_out_xl + j) * a.n_t;

  if (kj < 0 || kj >= a.K) {  // migration_kernel.cpp:79-85
    if (WRITES_OUT && !a.accumulate && active) orow[s] = 0.0f;
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
  LaneParams p{0.0f, 0.0f, -1.0f, 1.0f, 0, 0, 0.0f};
  bool needs_window = false;
  if (src >= 0) {
    // velocity scale: travel time and weight use alpha v, the aperture mask
    // and window alpha_mask v (both 1 for the reference operator, where the
    // products are exact)
    const float vt = a.alpha * v, vm = a.alpha_mask * v;
    p.t0sq = t0 * t0;
    const float inv_v2 = 1.0f / (vt * vt + 1e-10f);
    p.a = 4.0f * inv_v2;
    const float vth = vm * t0 * 0.5f;
    p.R2 = vth * vth * a.tan2;
    const int w = (int)(sqrt_rn(p.R2) * a.inv_dy) + 1;
    p.wlo = max(0, src - w);
    p.whi = min(a.K, src + w + 1);
    p.v2 = vt * vt;
    if constexpr (MODE == kAdjoint) p.mval = a.image[((size_t)io * a.n_out_xl + j) * a.n_t + s];
    if (src != kj) {
      // crosslines this lane could reach around kj, clipped to the data
      const int lo_need = max(a.xd_off, kj - w);
      const int hi_need = min(a.xd_off + a.n_xd, kj + w + 1);
      needs_window = (p.wlo > lo_need) || (p.whi < hi_need);
    }
  }
  // Sweep radius.  Counting needs every pair inside the mask.  The other
  // modes also stop where the travel time has passed the end of the trace in
  // every lane: with q = tau/dt >= n_t - 1 a term is zero (:298), and q is
  // nondecreasing in r^2, so the lane's last useful r^2 is below
  //   r2_end = ((n_t - 1/2) dt)^2 - t0^2) / a
  // (half a sample of slack, evaluated in double and rounded up, so no term
  // that the float arithmetic gathers is cut off).
  float L = p.R2;
  if constexpr (MODE != kCount) {
    if (src >= 0) {
      const double t_end = ((double)(a.n_t - 1) + 0.5) / (double)a.inv_dt;
      const double r2_end = (t_end * t_end - (double)p.t0sq) / (double)p.a;
      L = fminf(L, (float)(r2_end * (1.0 + 1e-6)));
    }
  }
  const float R2max = uniform_f(wave_max(L));
  float acc = 0.0f;
  unsigned cap = 0, cg = 0;
  if (R2max >= 0.0f) {
    if (__any(needs_window))
      sum_aperture<ST, GENERIC_ROWS, true, MODE>(a, tk, kj, p, R2max, acc, cap, cg);
    else
      sum_aperture<ST, GENERIC_ROWS, false, MODE>(a, tk, kj, p, R2max, acc, cap, cg);
  }
  if constexpr (COUNT) {
    if (active && a.out) orow[s] = (float)cap;
    if (active && a.out_g) a.out_g[((size_t)io * a.n_out_xl + j) * a.n_t + s] = (float)cg;
    const unsigned tap = wave_sum_u32(cap), tg = wave_sum_u32(cg);
    if (lane == 0) {
      atomicAdd(&a.counters[0], (unsigned long long)tap);
      atomicAdd(&a.counters[1], (unsigned long long)tg);
    }
  } else if (WRITES_OUT && active) {
    if (a.accumulate) {
      if (src >= 0) orow[s] += acc;
    } else {
      orow[s] = (src >= 0) ? acc : 0.0f;  // :148-151, :375
    }
  }
}

template <int ST, int WPB, bool GENERIC_ROWS, int MODE>
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
  migrate_trace_tile<ST, GENERIC_ROWS, MODE>(a, tile, io, j);
}

// Persistent batch kernel: every workgroup takes work items (one
// (volume, t0 tile, crossline group) each) from a global counter until the
// queue is drained, so volumes of very different size and cost share one
// launch.  Items are ordered by segment: segment k = the groups of one t0
// tile of one volume, latest tiles of all volumes first.  Each workgroup
// exits once the counter passes `total`.
template <int ST, int WPB>
__global__ __launch_bounds__(kWave * WPB) void batch_kernel(const MigArgs* __restrict__ items,
                                                            const BatchSeg* __restrict__ segs,
                                                            int n_segs,
                                                            unsigned* __restrict__ counter,
                                                            unsigned total) {
  __shared__ unsigned s_next;
  for (;;) {
    if (threadIdx.x == 0) s_next = atomicAdd(counter, 1u);
    __syncthreads();
    const unsigned w = s_next;
    __syncthreads();
    if (w >= total) return;
    int lo = 0, hi = n_segs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if ((unsigned)segs[mid].first <= w) lo = mid; else hi = mid - 1;
    }
    const BatchSeg sg = segs[lo];
    const int spatial = (int)w - sg.first;
    const MigArgs& a = items[sg.item];
    const int io = spatial / sg.n_jgroups;
    const int j = (spatial % sg.n_jgroups) * WPB + uniform_i((int)(threadIdx.x / kWave));
    if (io < a.n_out_il && j < a.n_out_xl) migrate_trace_tile<ST, false, kForward>(a, sg.tile, io, j);
  }
}

// FP64 dot product of two float32 vectors in two deterministic passes: block
// partial sums (fixed grid, fixed tree order), then one block over the
// partials.
constexpr int kDotBlocks = 1024;

__global__ __launch_bounds__(256) void dot_partial_kernel(const float* __restrict__ x,
                                                          const float* __restrict__ y, size_t n,
                                                          double* __restrict__ partial) {
  __shared__ double sh[256];
  double s = 0.0;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    s = __builtin_fma((double)x[i], (double)y[i], s);
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void dot_final_kernel(const double* __restrict__ partial, int n,
                                                        double* __restrict__ out) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += partial[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

__global__ void fill_zero_kernel(float* __restrict__ p, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0.0f;
}

// ---------------------------------------------------------------- packing
// Pair layout [n_rows + 1][n_xd + 1][n_t]: sample n_t-1 of every trace, the
// last trace of every row and the whole last row are zero.  F32 stores
// float2 (x[i], x[i+1] - x[i]).  F16 / BF16 store, per trace with
// s = max|x| (1 if the trace is all zero), the pair (x[i]/s, (x[i+1]-x[i])/s)
// rounded to nearest in the half format (low half = sample, high half =
// difference), and scale[row][xl] = s (0 for the zero trace and row).
// One wave per trace.
template <int ST>
__global__ __launch_bounds__(256) void pack_kernel(const float* __restrict__ x, void* __restrict__ xp,
                                                   float* __restrict__ scale, int n_rows,
                                                   int n_xd, int n_t) {
  const long long wave = (long long)blockIdx.x * 4 + threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const long long n_tr = (long long)(n_rows + 1) * (n_xd + 1);
  if (wave >= n_tr) return;
  const int row = (int)(wave / (n_xd + 1)), xl = (int)(wave % (n_xd + 1));
  const bool real = row < n_rows && xl < n_xd;
  const float* src = real ? x + ((size_t)row * n_xd + xl) * n_t : nullptr;
  float s = 1.0f;
  if constexpr (ST != PSM_STORE_F32) {
    float m = 0.0f;
    if (real)
      for (int i = lane; i < n_t; i += kWave) m = fmaxf(m, fabsf(src[i]));
    m = wave_max(m);
    s = (m > 0.0f) ? m : 1.0f;
    if (lane == 0) scale[wave] = real ? s : 0.0f;
  }
  const float inv = 1.0f / s;
  for (int i = lane; i < n_t; i += kWave) {
    float lo = 0.0f, d = 0.0f;
    if (real && i < n_t - 1) {
      lo = src[i];
      d = src[i + 1] - lo;
    }
    const size_t o = (size_t)wave * n_t + i;
    if constexpr (ST == PSM_STORE_F32) {
      static_cast<float2*>(xp)[o] = make_float2(lo, d);
    } else if constexpr (ST == PSM_STORE_F16) {
      const _Float16 hl = (_Float16)(lo * inv), hd = (_Float16)(d * inv);
      static_cast<unsigned*>(xp)[o] = (unsigned)__builtin_bit_cast(unsigned short, hl) |
                                      ((unsigned)__builtin_bit_cast(unsigned short, hd) << 16);
    } else {
      // round to nearest even into bfloat16
      auto bf = [](float v) -> unsigned {
        const unsigned u = __float_as_uint(v);
        return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
      };
      static_cast<unsigned*>(xp)[o] = bf(lo * inv) | (bf(d * inv) << 16);
    }
  }
}

// Table velocity provider (TableVelocityProvider::interpolateVelocity,
// velocity_reader.cpp:320-361): one thread per output sample, points in
// order, the 4-D distance in double, float weights and sums; the first point
// within 1e-10 gives the value directly.  Every thread reads the same point
// at a time, so the point loads are uniform.
__global__ __launch_bounds__(256) void table_velocity_kernel(
    const psm_velocity_point* __restrict__ pts, int n_pts, float vmin,
    const int32_t* __restrict__ il, const int32_t* __restrict__ xl, int n_il, int n_xl, int n_t,
    double dt, float* __restrict__ vel) {
  const size_t n = (size_t)n_il * n_xl * n_t;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += stride) {
    const int t = (int)(idx % n_t);
    const size_t r = idx / n_t;
    const int32_t xv = xl[r % n_xl], iv = il[r / n_xl];
    const double time = (double)t * dt;
    float sv = 0.0f, sw = 0.0f, out = 0.0f;
    bool exact = false;
    for (int k = 0; k < n_pts; ++k) {
      const psm_velocity_point p = pts[k];
      const double di = fabs((double)(p.inline_label - iv));
      const double dx = fabs((double)(p.crossline_label - xv));
      const double dtt = fabs(p.time - time);
      const double d = sqrt(di * di + dx * dx + dtt * dtt);
      if (d < 1e-10) {
        out = p.velocity;
        exact = true;
        break;
      }
      const float w = 1.0f / (1.0f + (float)d);
      sv = sv + p.velocity * w;
      sw = sw + w;
    }
    if (!exact) {
      out = sw > 0.0f ? sv / sw : vmin;
      if (out <= 0.0f) out = vmin;
    }
    vel[idx] = out;
  }
}

}  // namespace

size_t pairs_bytes(int storage, int n_rows, int n_xd, int n_t) {
  const size_t per = storage == PSM_STORE_F32 ? 8 : 4;
  return (size_t)(n_rows + 1) * (n_xd + 1) * n_t * per;
}

hipError_t launch_pack(int storage, const float* x, void* xp, float* scale, int n_rows, int n_xd,
                       int n_t, hipStream_t stream) {
  const long long n_tr = (long long)(n_rows + 1) * (n_xd + 1);
  const long long blocks = (n_tr + 3) / 4;
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  switch (storage) {
    case PSM_STORE_F16:
      hipLaunchKernelGGL(pack_kernel<PSM_STORE_F16>, dim3((unsigned)blocks), dim3(256), 0, stream,
                         x, xp, scale, n_rows, n_xd, n_t);
      break;
    case PSM_STORE_BF16:
      hipLaunchKernelGGL(pack_kernel<PSM_STORE_BF16>, dim3((unsigned)blocks), dim3(256), 0, stream,
                         x, xp, scale, n_rows, n_xd, n_t);
      break;
    default:
      hipLaunchKernelGGL(pack_kernel<PSM_STORE_F32>, dim3((unsigned)blocks), dim3(256), 0, stream,
                         x, xp, scale, n_rows, n_xd, n_t);
      break;
  }
  return hipGetLastError();
}

template <int ST, int WPB, int MODE>
hipError_t launch_wpb(const MigArgs& a, hipStream_t stream) {
  const int n_tiles = (a.n_t + kWave - 1) / kWave;
  const int n_jgroups = (a.n_out_xl + WPB - 1) / WPB;
  const long long per_tile = (long long)a.n_out_il * n_jgroups;
  const long long per_tile_pad = (per_tile + kXcds - 1) / kXcds * kXcds;
  const long long blocks = per_tile_pad * n_tiles;
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  const dim3 grid((unsigned)blocks), block(kWave * WPB);
  const bool generic = a.row_pos != nullptr;
  const int ptp = (int)per_tile_pad;
  if (generic)
    hipLaunchKernelGGL((migrate_kernel<ST, WPB, true, MODE>), grid, block, 0, stream, a, n_tiles, n_jgroups, ptp);
  else
    hipLaunchKernelGGL((migrate_kernel<ST, WPB, false, MODE>), grid, block, 0, stream, a, n_tiles, n_jgroups, ptp);
  return hipGetLastError();
}

template <int ST, int MODE>
hipError_t launch_st(const MigArgs& a, hipStream_t stream, int wpb) {
  switch (wpb) {
    case 1: return launch_wpb<ST, 1, MODE>(a, stream);
    case 2: return launch_wpb<ST, 2, MODE>(a, stream);
    case 8: return launch_wpb<ST, 8, MODE>(a, stream);
    case 16: return launch_wpb<ST, 16, MODE>(a, stream);
    default: return launch_wpb<ST, 4, MODE>(a, stream);
  }
}

hipError_t launch_migrate(const MigArgs& a0, int mode, hipStream_t stream, const Tuning& tune,
                          int row_extent) {
  if (a0.n_out_il <= 0 || a0.n_out_xl <= 0 || a0.n_t <= 0) return hipSuccess;
  MigArgs a = a0;
  // aperture-row passes: pass p sums the row offsets [p*R, (p+1)*R) (generic
  // rows: row indices); passes after the first accumulate into `out`
  const bool f32_traces = mode != kForward || a0.storage == PSM_STORE_F32;
  const int rpp = tune.rows_per_pass >= 0 ? tune.rows_per_pass : (f32_traces ? kAutoRowsF32 : 0);
  const int R = (mode == kCount || rpp <= 0) ? (row_extent + 1) : rpp;
  const int wpb = tune.waves_per_block;
  for (int lo = 0, pass = 0; lo <= row_extent; lo += R, ++pass) {
    a.d_lo = lo;
    a.d_hi = lo + R;
    a.accumulate = pass > 0 ? 1 : 0;
    hipError_t e;
    switch (mode) {
      case kCount: e = launch_st<PSM_STORE_F32, kCount>(a, stream, wpb); break;
      case kAdjoint: e = launch_st<PSM_STORE_F32, kAdjoint>(a, stream, wpb); break;
      case kTangent: e = launch_st<PSM_STORE_F32, kTangent>(a, stream, wpb); break;
      default:
        switch (a.storage) {
          case PSM_STORE_F16: e = launch_st<PSM_STORE_F16, kForward>(a, stream, wpb); break;
          case PSM_STORE_BF16: e = launch_st<PSM_STORE_BF16, kForward>(a, stream, wpb); break;
          default: e = launch_st<PSM_STORE_F32, kForward>(a, stream, wpb); break;
        }
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_dot_f64(const float* x, const float* y, size_t n, double* partial,
                          double* out, hipStream_t stream) {
  hipLaunchKernelGGL(dot_partial_kernel, dim3(kDotBlocks), dim3(256), 0, stream, x, y, n, partial);
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(256), 0, stream, partial, kDotBlocks, out);
  return hipGetLastError();
}

size_t dot_scratch_doubles() { return kDotBlocks + 1; }

template <int ST>
hipError_t launch_batch_st(const MigArgs* items, const BatchSeg* segs, int n_segs, unsigned* counter,
                           unsigned total, int wpb, int blocks, hipStream_t stream) {
  const dim3 grid((unsigned)blocks);
  switch (wpb) {
    case 1: hipLaunchKernelGGL((batch_kernel<ST, 1>), grid, dim3(kWave), 0, stream, items, segs, n_segs, counter, total); break;
    case 2: hipLaunchKernelGGL((batch_kernel<ST, 2>), grid, dim3(2 * kWave), 0, stream, items, segs, n_segs, counter, total); break;
    case 8: hipLaunchKernelGGL((batch_kernel<ST, 8>), grid, dim3(8 * kWave), 0, stream, items, segs, n_segs, counter, total); break;
    default: hipLaunchKernelGGL((batch_kernel<ST, 4>), grid, dim3(4 * kWave), 0, stream, items, segs, n_segs, counter, total); break;
  }
  return hipGetLastError();
}

hipError_t launch_batch(int storage, const MigArgs* items, const BatchSeg* segs, int n_segs,
                        unsigned* counter, unsigned total, int wpb, int blocks, hipStream_t stream) {
  switch (storage) {
    case PSM_STORE_F16: return launch_batch_st<PSM_STORE_F16>(items, segs, n_segs, counter, total, wpb, blocks, stream);
    case PSM_STORE_BF16: return launch_batch_st<PSM_STORE_BF16>(items, segs, n_segs, counter, total, wpb, blocks, stream);
    default: return launch_batch_st<PSM_STORE_F32>(items, segs, n_segs, counter, total, wpb, blocks, stream);
  }
}

hipError_t launch_table_velocity(const psm_velocity_point* pts, int n_pts, float vmin,
                                 const int32_t* il, const int32_t* xl, int n_il, int n_xl,
                                 int n_t, double dt, float* vel, hipStream_t stream) {
  const size_t n = (size_t)n_il * n_xl * n_t;
  if (n == 0) return hipSuccess;
  size_t blocks = (n + 255) / 256;
  if (blocks > 262144) blocks = 262144;
  hipLaunchKernelGGL(table_velocity_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, pts, n_pts,
                     vmin, il, xl, n_il, n_xl, n_t, dt, vel);
  return hipGetLastError();
}

hipError_t launch_fill_zero(float* p, size_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, p, n);
  return hipGetLastError();
}

}  // namespace psm
