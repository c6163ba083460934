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
// -ffast-math build contracts them), q = tau * (1/dt) with 1/dt rounded once.
// tau is the correctly rounded square root (the summing modes take it and
// the weight from one reciprocal square root and a Newton step, the pair
// count from a rounded root and a neighbour test).  The file is compiled with
// -ffp-contract=off so that these FMAs are exactly the ones written here.
#include <hip/hip_runtime.h>

#include <cstddef>
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

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, kWave));
  return v;
}

// Largest v over lanes 0 .. N-1 (N a power of two, at most 16), as a
// scalar: DPP row_shr:1, 2, 4, 8 inside the first row of 16 lanes (a lane
// whose source lies before its row keeps its own value), then lane N-1.
// Called with every lane active.
template <int N>
__device__ __forceinline__ int lanes_max_i(int v) {
  static_assert(N >= 1 && N <= 16 && (N & (N - 1)) == 0, "N lanes of one DPP row");
  if constexpr (N >= 2) v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x111, 0xf, 0xf, false));
  if constexpr (N >= 4) v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x112, 0xf, 0xf, false));
  if constexpr (N >= 8) v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x114, 0xf, 0xf, false));
  if constexpr (N >= 16) v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x118, 0xf, 0xf, false));
  return __builtin_amdgcn_readlane(v, N - 1);
}

__device__ __forceinline__ unsigned wave_sum_u32(unsigned v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Lane l receives v of lane l - 1 (lane 0 receives 0): DPP wave_shr:1.
__device__ __forceinline__ int lane_prev(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, true);
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
// A whole struct read that way (dword by dword: the implicit copy does not
// bind across address spaces).
template <class T>
__device__ __forceinline__ T const_load(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "dword-sized struct");
  using CU = const __attribute__((address_space(4))) unsigned*;
  T v;
  unsigned* d = reinterpret_cast<unsigned*>(&v);
  const CU q = (CU)p;
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / 4); ++k) d[k] = q[k];
  return v;
}

// ---------------------------------------------------------------- storage
// One stored pair (x[i], x[i+1] - x[i]) and its linear interpolation in fp32.
template <int ST>
struct Store;

template <>
struct Store<PSM_STORE_F32> {
  using Raw = float2;
  static constexpr int kBytes = 8;
  static constexpr int kShift = 3;  // log2 kBytes
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
  float t0sq, a, R2;
  float iv2;  // 1 / (v^2 + 1e-10) = a / 4
  int wlo, whi;
  float mval;  // adjoint: the image sample this lane scatters
};

// Travel time and amplitude weight of one pair from one reciprocal square
// root (migration_kernel.cpp:212, 270): with y = t0^2 + a r^2 and
// rs = rsq(y) (v_rsq_f32, within 1 ulp), s = y rs is sqrt(y) to about
// 1.5 ulp, and one Newton step tau = s + (y - s^2) rs / 2 (the residual
// from an FMA) gives the correctly rounded root but for arguments within a
// tiny fraction of an ulp of a rounding midpoint -- the reference rounds tau
// correctly, and the bound i < n_t - 1 makes terms at the trace end depend
// on the last bit of tau.  The weight 1 / (tau v^2 + 1e-10) is taken as
// rs / (v^2 + 1e-10) (the two differ by a relative 1e-10 (tau - 1) /
// (tau v^2), nothing at seismic velocities).  The forward, adjoint and
// tangent all use this one form, so the adjoint is the exact transpose of
// the forward.
struct TauW {
  float tau, w;
};
__device__ __forceinline__ TauW tau_weight(const LaneParams& p, float r2, bool amp) {
  const float y = __builtin_fmaf(p.a, r2, p.t0sq);
  const float rs = __builtin_amdgcn_rsqf(y);
  const float s = y * rs;
  const float e = __builtin_fmaf(-s, s, y);
  return TauW{__builtin_fmaf(e, 0.5f * rs, s), amp ? rs * p.iv2 : 1.0f};
}

// Sums one aperture row pair of one (output trace, t0 tile): rows +db and
// -db of the output inline (rowp / rowm, present when hp / hm; ip / im their
// index in the scale table), d2 = (db dx)^2, over the crossline offsets with
// r^2 <= R2max.  WINDOW = some lane uses parameters of an earlier crossline
// whose crossline window can cut the aperture (sticky samples of
// migration_kernel.cpp:64-152), so each gather is also checked against that
// window.  MODE: kForward sums, kCount counts pairs, kAdjoint scatters the
// lane's image sample into the data cube with the same (i, f, w, mask) as the
// forward sum, kTangent sums the derivative of each term with respect to the
// velocity scale.  All loop bounds and trace offsets are wave-uniform (scalar
// registers); absent crosslines read the row's zero trace and absent rows the
// zero row, so the four gathers of a step are issued back to back, and the
// fast path loads step dk+1 before it consumes step dk (two steps of gathers
// in flight).
template <int ST, bool WINDOW, int MODE>
__device__ __forceinline__ void sweep_row(const MigArgs& a, int kj, const LaneParams& p,
                                          float R2max, float& acc, unsigned& cap, unsigned& cg,
                                          const char* rowp, bool hp, int ip, const char* rowm,
                                          bool hm, int im, float d2) {
  using S = Store<ST>;
  constexpr bool COUNT = MODE == kCount;
  const int n_t = a.n_t;
  const int nt1 = n_t - 1;
  const int xd_lo = a.xd_off, xd_hi = a.xd_off + a.n_xd;
  // pair layout rows are [n_xd + 1][n_t]: the last trace of every row is zero
  const int tb = n_t * S::kBytes;
  const size_t row_bytes = (size_t)(a.n_xd + 1) * tb;
  const int row_tr = a.n_xd + 1;  // traces per row (scales)
  // largest crossline offset that can hold data
  const int dk_lim = max(kj - xd_lo, xd_hi - 1 - kj);

  // last dk with r^2 <= R2max (r^2 is nondecreasing in dk)
  // (v_sqrt_f32 is within one ulp; the loop finds the last offset)
  int kmax = (int)(__builtin_amdgcn_sqrtf(R2max - d2) * a.inv_dy) + 2;
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
        const TauW tw = tau_weight(p, r2, a.amp);
        const float q = tw.tau * a.inv_dt;
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
        g.w = tw.w;
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
      // (the count evaluates i from the correctly rounded root)
      const TauW tw = COUNT ? TauW{sqrt_rn(__builtin_fmaf(p.a, r2, p.t0sq)), 1.0f}
                            : tau_weight(p, r2, a.amp);
      const float tau = tw.tau;
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
          const float wm = tw.w * p.mval;
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
        const float w = tw.w;
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
}

// Sums the aperture of one (output trace, t0 tile) over the row offsets
// |db| in [d_lo, d_hi) of this pass (generic rows: the data rows
// [d_lo, d_hi)), one sweep_row per row pair.
template <int ST, bool GENERIC_ROWS, bool WINDOW, int MODE>
__device__ __forceinline__ void sum_aperture(const MigArgs& a, const Task& tk, int kj,
                                             const LaneParams& p, float R2max,
                                             float& acc, unsigned& cap, unsigned& cg) {
  using S = Store<ST>;
  const int tb = a.n_t * S::kBytes;
  const size_t row_bytes = (size_t)(a.n_xd + 1) * tb;
  const char* const base = reinterpret_cast<const char*>(a.xp);
  const char* const zero_row = base + (size_t)a.n_rows * row_bytes;
  auto row_sweep = [&](const char* rowp, bool hp, int ip, const char* rowm, bool hm, int im,
                       float d2) {
    sweep_row<ST, WINDOW, MODE>(a, kj, p, R2max, acc, cap, cg, rowp, hp, ip, rowm, hm, im, d2);
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

// Image-point parameters of one lane (sample s of output trace (io, j),
// buffer crossline kj in [0, K)) and its sweep radius.
struct LaneSetup {
  LaneParams p;
  int src;            // crossline whose velocity the lane uses; -1: the sample is not valid
  bool needs_window;  // src != kj and its crossline window can cut the aperture
  float L;            // last useful r^2 of this lane (-1: none)
};

template <int MODE>
__device__ __forceinline__ LaneSetup lane_setup(const MigArgs& a, const Task& tk, int io, int j,
                                                int kj, int s, bool active) {
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
  // (a lane without a sample keeps t0^2 = 1, so its reciprocal square root
  // stays finite; its terms are masked or discarded)
  LaneParams p{1.0f, 0.0f, -1.0f, 0.0f, 0, 0, 0.0f};
  bool needs_window = false;
  if (src >= 0) {
    // velocity scale: travel time and weight use alpha v, the aperture mask
    // and window alpha_mask v (both 1 for the reference operator, where the
    // products are exact)
    const float vt = a.alpha * v, vm = a.alpha_mask * v;
    p.t0sq = t0 * t0;
    const float inv_v2 = 1.0f / (vt * vt + 1e-10f);
    p.a = 4.0f * inv_v2;
    p.iv2 = inv_v2;
    const float vth = vm * t0 * 0.5f;
    p.R2 = vth * vth * a.tan2;
    const int w = (int)(sqrt_rn(p.R2) * a.inv_dy) + 1;
    p.wlo = max(0, src - w);
    p.whi = min(a.K, src + w + 1);
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
  return LaneSetup{p, src, needs_window, L};
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
  const LaneSetup ls = lane_setup<MODE>(a, tk, io, j, kj, s, active);
  const LaneParams& p = ls.p;
  const int src = ls.src;
  const bool needs_window = ls.needs_window;
  const float L = ls.L;
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

// ---------------------------------------------------------------- LDS-staged forward path
// A workgroup of W waves migrates the W adjacent output crosslines
// j0 .. j0+W-1 of one output inline over one t0 tile.  For each aperture row
// pair (+db, -db) the crossline offsets dk are swept in groups of G steps.
// Wave w at step dk reads the columns kj_w + dk ("side 0") and kj_w - dk
// ("side 1"), i.e. the column offsets o = w + dk on side 0 (column j0 + o)
// and o = W-1-w + dk on side 1 (column j0 + W-1 - o): all waves walk the same
// two column sequences, and each column is read by up to W waves at W
// consecutive steps.  The stored pairs any of those reads can touch lie in a
// window [lo, hi] bounded from the waves' smallest and largest t0^2 and 4/v^2
// (q = sqrt(t0^2 + a r^2)/dt is nondecreasing in both); the window is copied
// once into LDS (16 bytes per lane, one coalesced load per window) into a ring
// of slots per (row, side) that is filled one group ahead of the reads.  The
// waves gather from LDS instead of through the texture path and keep the
// four-partner sharing of tau, i, f and the weight; a column outside the data
// is staged as zeros.  A row whose windows exceed a slot, or whose column
// offsets exceed the column table, runs the gather path (sweep_row); so does
// a workgroup with lanes whose sticky crossline window can cut the aperture.
constexpr int kWinPairs = 128;  // stored pairs per staged window
constexpr int kMaxCols = 512;   // column offsets per side one row can stage

// One staging job (column offset o, side, ring row rr; ok = the job exists)
// and the table values of one group of steps (r^2 and the two column
// addresses per step, the column table entries of the staging jobs).
struct StageJob {
  int o, side, rr;
  bool ok;
};
template <int G, int J>
struct StageTab {
  float r2[G];
  int bR[G], bL[G];
  int mj[J], mw[J];
};

// LDS of one staged workgroup.  Forward: slots of kWinPairs stored pairs.
// Adjoint: slots of kWinPairs accumulator pairs (A, B) -- the window's
// samples lo .. lo + kWinPairs - 1 receive d[j] = A[j] - B[j] + B[j - 1] --
// in G copies per column, and a table of the samples each column flushes.
template <int ST, int W, int G, int MODE = kForward>
struct LdsPlan {
  static constexpr bool kAdj = MODE == kAdjoint;
  // Forward: the rows +db and -db of a row pair share every step's tau, i,
  // f, weight and mask, so a column's windows of the two rows are staged as
  // one window of summed FP32 pairs (x_p[i] + x_m[i], d_p[i] + d_m[i]) --
  // one ring row, half the LDS reads and interpolations per step.  FP16 /
  // BF16 rows are widened and multiplied by their trace's scale as they are
  // staged (s_p x_p[i] + s_m x_m[i], ...), so the steps are the FP32 steps.
  // The adjoint likewise adds the same terms to both rows, so one
  // accumulator window per column is flushed into both rows.
  static constexpr bool kSumRows = !kAdj;
  static constexpr int kRingRows = 1;
  // forward: a window may run past the trace end -- its pairs past n_t - 1
  // are zeros in the slot -- so the steps read pair i without clamping it
  // (two pairs per staging lane: 16 bytes of FP32 pairs, 8 of half pairs)
  static constexpr bool kOpenEnd = kSumRows;
  static constexpr int kB = 8;                      // LDS bytes per pair (FP32 pair or (A, B))
  static constexpr int kHB = Store<ST>::kBytes;     // bytes per stored pair in the cube
  static constexpr int kHShift = kHB == 8 ? 0 : 1;  // LDS bytes -> cube bytes
  static constexpr int kSlot = kWinPairs * kB;      // bytes of one window
  // adjoint: copy u of a column's window takes the adds of step u of a group
  // (a column is touched by a different wave at each step, so waves that
  // drift apart within a group never share a copy)
  static constexpr int kCopies = kAdj ? G : 1;
  // column offsets per side one row can stage (the adjoint's larger windows
  // leave room for fewer table entries)
  static constexpr int kMaxColsP = kAdj ? kMaxCols / 2 : kMaxCols;
  static constexpr int kCols = kMaxColsP + 4 * G + W;    // table entries (reads past nO stay inside)
  // bytes of a window-table entry: forward (cube bytes of the window's
  // first pair, LDS bytes of its pairs lo .. hi), adjoint (samples flushed)
  static constexpr int kWidB = kAdj ? 4 : 8;
  static constexpr int kTables = 3 * kCols * 4 + 2 * kCols * kWidB + 5 * W * 4 + W * 4;
  // Double ring (2G + W - 1 slots per ring row and side): the next group's
  // windows are stored while the current group is read, one barrier per
  // group.  Single ring (G + W - 1 slots): they are stored after a barrier
  // that ends the current group's reads, two barriers per group; used when
  // the double ring would leave fewer than three workgroups per CU.
  static constexpr bool kSingle =
      !kAdj && (2 * kRingRows * (2 * G + W - 1) + 1) * kSlot + kTables > 53 * 1024;
  static constexpr int kRing = kSingle ? G + W - 1 : 2 * G + W - 1;
  static constexpr int kSide = kRing * kSlot * kCopies;  // side 0 -> side 1
  static constexpr int kRow = 2 * kSide;                 // ring row 0 -> ring row 1
  // one slot that absent staging jobs write (forward)
  static constexpr int kSpare = kRingRows * kRow;
  static constexpr int kMeta = kSpare + (kAdj ? 0 : kSlot);  // int meta[2][kCols]
  static constexpr int kR2 = kMeta + 2 * kCols * 4;      // float r2tab[kCols]
  static constexpr int kPar = kR2 + kCols * 4;           // float par[5][W]
  static constexpr int kFlag = kPar + 5 * W * 4;         // int flag[W]
  // wid[2][kCols]: forward, int2 (lo * kHB: the cube byte offset of the
  // window's first pair in its trace; bytes of the window's pairs lo .. hi
  // in LDS); adjoint, int: samples the column flushes
  static constexpr int kWid = kFlag + W * 4;
  static constexpr int kBytes = kWid + 2 * kCols * kWidB;
  // waves per SIMD the LDS allows (three 8-wave workgroups per CU with the
  // double ring), else no register target
  static constexpr int kWavesPerEu = (W == 8 && !kSingle && 3 * (kBytes + 16) <= 160 * 1024) ? 6 : 1;
  static_assert(kRingRows == 1 || kRow <= 65535, "ring row offset must fit a ds_read offset");
  static_assert(kAdj || kSlot <= 16 * kWave, "one load per window");
  static_assert(!kAdj || !kSingle, "the adjoint uses the double ring");
  static_assert(!kAdj || 3 * (kBytes + 16) <= 160 * 1024, "three adjoint workgroups per CU");
  __device__ static int slot(int side, int o) { return side * kSide + (o % kRing) * kCopies * kSlot; }
};


// One workgroup of the staged form: output crosslines j0 .. j0 + W - 1 of
// output inline io over t0 tile `tile`; smem holds LdsPlan<ST, W, G,
// MODE>::kBytes of LDS.  Every wave of the workgroup calls it with the same
// arguments.  MODE = kForward sums the staged windows into the output;
// kAdjoint scatters the image samples into accumulator windows that are
// flushed into the data cube (ST = F32).
template <int ST, int W, int G, bool AMP, int MODE = kForward>
__device__ __forceinline__ void staged_tile(const MigArgs a, const Task tk, int tile, int io,
                                            int j0, char* smem) {
  using S = Store<ST>;
  using PL = LdsPlan<ST, W, G, MODE>;
  constexpr int kB = PL::kB;
  const int w = uniform_i((int)(threadIdx.x / kWave));
  const int lane = threadIdx.x % kWave;
  const int j = j0 + w;
  const int kj = a.cur_xl_start + j;
  const int s = tile * kWave + lane;
  const bool active = s < a.n_t;
  const bool in_out = j < a.n_out_xl;
  const bool in_buf = kj >= 0 && kj < a.K;

  LaneSetup ls{{1.0f, 0.0f, -1.0f, 0.0f, 0, 0, 0.0f}, -1, false, -1.0f};
  if (in_out && in_buf) ls = lane_setup<MODE>(a, tk, io, j, kj, s, active);
  const bool valid = ls.src >= 0;
  const float R2w = uniform_f(wave_max(ls.L));
  // smallest aperture R^2 of the wave's valid lanes (a lane without a valid
  // sample needs no mask: its sum is discarded)
  const float R2in = uniform_f(wave_min(valid ? ls.p.R2 : 3.0e38f));
  const float Tlo2 = uniform_f(wave_min(valid ? ls.p.t0sq : 3.0e38f));
  const float Thi2 = uniform_f(wave_max(valid ? ls.p.t0sq : 0.0f));
  const float Alo = uniform_f(wave_min(valid ? ls.p.a : 3.0e38f));
  const float Ahi = uniform_f(wave_max(valid ? ls.p.a : 0.0f));
  const bool win = __any(ls.needs_window);

  int* const meta = reinterpret_cast<int*>(smem + PL::kMeta);
  float* const r2tab = reinterpret_cast<float*>(smem + PL::kR2);
  float* const par = reinterpret_cast<float*>(smem + PL::kPar);
  int* const flag = reinterpret_cast<int*>(smem + PL::kFlag);
  if (lane == 0) {
    par[0 * W + w] = R2w;
    par[1 * W + w] = Tlo2;
    par[2 * W + w] = Thi2;
    par[3 * W + w] = Alo;
    par[4 * W + w] = Ahi;
    flag[w] = win ? 1 : 0;
  }
  __syncthreads();
  // lane l < W holds the parameters of wave l
  const bool pl = lane < W;
  const float pR2 = pl ? par[lane] : -1.0f;
  const float pTlo2 = pl ? par[W + lane] : 0.0f;
  const float pThi2 = pl ? par[2 * W + lane] : 0.0f;
  const float pAlo = pl ? par[3 * W + lane] : 0.0f;
  const float pAhi = pl ? par[4 * W + lane] : 0.0f;
  const bool any_win = __any(pl && flag[pl ? lane : 0] != 0);
  const float R2wg = uniform_f(wave_max(pR2));
  // workgroup bounds of t0^2 and 4/v^2 over the valid lanes of all waves
  const float gTlo2 = uniform_f(wave_min(pl ? pTlo2 : 3.0e38f));
  const float gThi2 = uniform_f(wave_max(pl ? pThi2 : 0.0f));
  const float gAlo = uniform_f(wave_min(pl ? pAlo : 3.0e38f));
  const float gAhi = uniform_f(wave_max(pl ? pAhi : 0.0f));

  float acc = 0.0f;
  unsigned cap = 0, cg = 0;
  if (any_win) {
    // sticky crossline windows in the group: every wave runs the gather path
    if (R2w >= 0.0f) {
      if (win)
        sum_aperture<ST, false, true, MODE>(a, tk, kj, ls.p, R2w, acc, cap, cg);
      else
        sum_aperture<ST, false, false, MODE>(a, tk, kj, ls.p, R2w, acc, cap, cg);
    }
  } else if (R2wg >= 0.0f) {
    const int c = tk.c;
    const int nt1 = a.n_t - 1;
    const int xd_lo = a.xd_off, xd_hi = a.xd_off + a.n_xd;
    const int tb = a.n_t * PL::kHB;  // bytes of one stored trace
    const size_t row_bytes = (size_t)(a.n_xd + 1) * tb;
    const char* const base = reinterpret_cast<const char*>(a.xp);
    const char* const zero_row = base + (size_t)a.n_rows * row_bytes;
    const int row_tr = a.n_xd + 1;
    const int kbase = a.cur_xl_start + j0;  // buffer crossline of side-0 column offset 0
    // A lane without a valid sample reads at the lower corner of the
    // workgroup's bounds, inside every window; its terms are masked (R2 = -1).
    LaneParams p = ls.p;
    if (!valid) {
      p.t0sq = gTlo2;
      p.a = gAlo;
    }
    // Forward with amplitude correction: every term's weight is
    // rsq(tau^2) * (1/v^2) with the lane's one 1/v^2, so the sweep (staged
    // groups and gather rows alike) accumulates sum * rsq(tau^2) and the
    // lane's 1/v^2 multiplies the total once, at the end.
    const float lane_iv2 = p.iv2;
    if constexpr (MODE == kForward && AMP) p.iv2 = 1.0f;

    if constexpr (PL::kAdj) {
      // the accumulator ring starts at zero; every flush leaves its slot zero
      using Z4 = unsigned __attribute__((ext_vector_type(4)));
      for (int off = (int)threadIdx.x * 16; off < PL::kRingRows * PL::kRow; off += W * kWave * 16)
        *reinterpret_cast<Z4*>(smem + off) = Z4{0u, 0u, 0u, 0u};
    }

    int dmax = (int)(__builtin_sqrtf(R2wg) / a.dx) + 2;
    dmax = min(dmax, max(tk.b1 - 1 - c, c - tk.b0));
    dmax = uniform_i(min(dmax, a.d_hi - 1));
    // true while no wave reads LDS of the previous row: a staged forward row
    // ends with a barrier after its last reads (the parameter vote above and
    // the adjoint's flushes do not)
    bool quiet = false;
    // Last crossline offsets carried from row to row: the test r^2 =
    // fma(dk dy, dk dy, d^2) <= R^2 only gets harder as d grows, so each
    // row resumes from the previous row's answer and steps down.  km_run:
    // lane l < W, wave l's sweep (clipped to offsets that can hold data);
    // kin_run: every valid lane of this wave inside the aperture.
    // ---- staging helpers (forward)
    using V4 = unsigned __attribute__((ext_vector_type(4)));
    using V2 = unsigned __attribute__((ext_vector_type(2)));
    using F4 = float __attribute__((ext_vector_type(4)));
    using JV = std::conditional_t<S::kScaled, V2, V4>;  // one row of a job, per lane
    const int lb = lane * 16;
    auto load_row = [&](__amdgpu_buffer_rsrc_t r, int voff, int soff) -> JV {
      if constexpr (S::kScaled)
        return __builtin_bit_cast(V2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
      else
        return __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
    };
    // the two stored pairs of one row as FP32 pairs (sample, difference),
    // unscaled
    auto unpack = [&](JV v) -> F4 {
      if constexpr (ST == PSM_STORE_F32) {
        return __builtin_bit_cast(F4, v);
      } else if constexpr (ST == PSM_STORE_F16) {
        return F4{(float)__builtin_bit_cast(_Float16, (unsigned short)(v.x & 0xffffu)),
                  (float)__builtin_bit_cast(_Float16, (unsigned short)(v.x >> 16)),
                  (float)__builtin_bit_cast(_Float16, (unsigned short)(v.y & 0xffffu)),
                  (float)__builtin_bit_cast(_Float16, (unsigned short)(v.y >> 16))};
      } else {
        return F4{__uint_as_float(v.x << 16), __uint_as_float(v.x & 0xffff0000u),
                  __uint_as_float(v.y << 16), __uint_as_float(v.y & 0xffff0000u)};
      }
    };
    // one row's pairs times its trace's scale s (F16 / BF16; FP32 rows
    // are not scaled).  fma(x, s, 0) is x s but for the sign of an exact
    // zero (no matter in a sum), and lets an F16 x be widened and scaled
    // by one v_fma_mix_f32.
    auto widen = [&](JV v, float s) -> F4 {
      if constexpr (S::kScaled) {
        const F4 x = unpack(v);
        return F4{__builtin_fmaf(x.x, s, 0.0f), __builtin_fmaf(x.y, s, 0.0f),
                  __builtin_fmaf(x.z, s, 0.0f), __builtin_fmaf(x.w, s, 0.0f)};
      }
      return unpack(v);
    };
    // acc + the row's pairs (times s for F16 / BF16)
    auto widen_add = [&](JV v, float s, F4 acc) -> F4 {
      const F4 x = unpack(v);
      if constexpr (S::kScaled)
        return F4{__builtin_fmaf(x.x, s, acc.x), __builtin_fmaf(x.y, s, acc.y),
                  __builtin_fmaf(x.z, s, acc.z), __builtin_fmaf(x.w, s, acc.w)};
      return acc + x;
    };
    // A job's source: column o, side of the job index j within the
    // columns [oa, oa + nc).
    // (one return statement: a lambda returning this struct on two paths
    // made the compiler copy the kernel arguments to scratch)
    auto job_at = [&](int oa, int nc, int j) {
      return StageJob{oa + (j >> 1), j & 1, 0, j < 2 * nc};
    };
    int km_run = -1;
    if (pl && pR2 >= 0.0f) {
      const int kjl = kbase + lane;
      // (v_sqrt_f32 is within one ulp: the start is at or above the answer)
      km_run = min((int)(__builtin_amdgcn_sqrtf(pR2) * a.inv_dy) + 2,
                   max(kjl - xd_lo, xd_hi - 1 - kjl));
    }
    int kin_run = R2in >= 0.0f ? (int)(__builtin_amdgcn_sqrtf(R2in) * a.inv_dy) + 2 : -1;
    for (int d = max(0, a.d_lo); d <= dmax; ++d) {
      const float dxv = (float)d * a.dx;
      const float d2 = uniform_f(dxv * dxv);
      if (d2 > R2wg) break;
      const int rp = c + d, rm = c - d;
      const bool hp = rp >= tk.b0 && rp < tk.b1;
      const bool hm = d > 0 && rm >= tk.b0 && rm < tk.b1;
      if (!hp && !hm) continue;
      // last crossline offset of each wave's sweep (lane l < W: wave l), as
      // in sweep_row: r^2 <= R2max, clipped to offsets that can hold data
      if (pl) {
        if (pR2 < d2) km_run = -1;
        while (km_run >= 0) {
          const float dyv = (float)km_run * a.dy;
          if (__builtin_fmaf(dyv, dyv, d2) <= pR2) break;
          --km_run;
        }
      }
      const int kmax_wg = lanes_max_i<W>(km_run);
      if (kmax_wg < 0) continue;
      const int nO = kmax_wg + W;  // column offsets per side
      // last crossline offset inside every valid lane's aperture (r^2 as in
      // the column table), at most kmax_wg; -1 if none
      // (kmax_wg is nonincreasing in d as well, so the clip carries over)
      kin_run = R2in < d2 ? -1 : min(kin_run, kmax_wg);
      while (kin_run >= 0) {
        const float dyv = (float)kin_run * a.dy;
        if (__builtin_fmaf(dyv, dyv, d2) <= R2in) break;
        --kin_run;
      }
      const int kin = uniform_i(kin_run);
      const int r0 = hp ? rp : rm;  // ring row 0; ring row 1 = rm when both rows exist
      const bool two = hp && hm;

      // Window of column o of `side` in this row: stored pairs lo .. hi any
      // wave's read can touch (wbytes = LDS bytes of pairs lo .. min(hi,
      // n_t - 1); 0 if no wave reads the column); ovf |= the window exceeds
      // a slot.
      auto col_window = [&](int side, int o, int& lo, int& wbytes, bool& ovf) {
        const int dlo = max(side, o - (W - 1)), dhi = min(o, kmax_wg);
        lo = 0;
        wbytes = 0;
        if (dlo <= dhi) {
          const float ylo = (float)dlo * a.dy, yhi = (float)dhi * a.dy;
          const float lo_t = __builtin_amdgcn_sqrtf(
              __builtin_fmaf(gAlo, __builtin_fmaf(ylo, ylo, d2), gTlo2));
          const float hi_t = __builtin_amdgcn_sqrtf(
              __builtin_fmaf(gAhi, __builtin_fmaf(yhi, yhi, d2), gThi2));
          // one index of slack each way covers the rounding of q and the
          // one-ulp error of v_sqrt_f32 (a relative 1.2e-7 of q)
          if constexpr (PL::kOpenEnd) {
            // The window is not cut at the trace end; only pairs up to
            // n_t - 1 are loaded and the lanes past them load zeros.  The
            // lane holding pair n_t - 1 must not also hold pair n_t (the
            // next trace's first), so where the window reaches past the end
            // lo is lowered by one to make n_t - 1 - lo odd.
            lo = max((int)(lo_t * a.inv_dt) - 1, 0);
            const int hi = max((int)(hi_t * a.inv_dt) + 1, 0);
            if (hi > nt1 && ((nt1 - lo) & 1) == 0) {
              if (lo > 0)
                --lo;
              else
                ovf = true;
            }
            ovf = ovf || (hi - lo >= kWinPairs);
            wbytes = lo <= nt1 ? (min(hi, nt1) - lo + 1) * kB : 0;
          } else {
            lo = min(max((int)(lo_t * a.inv_dt) - 1, 0), nt1);
            const int hi = min(max((int)(hi_t * a.inv_dt) + 1, 0), nt1);
            // (the adjoint also touches sample hi + 1)
            ovf = ovf || (hi - lo >= (PL::kAdj ? kWinPairs - 1 : kWinPairs));
            wbytes = PL::kAdj ? min(hi + 1, nt1) - lo + 1 : (hi - lo + 1) * kB;
          }
        }
      };
      bool fallback = nO > PL::kMaxColsP;
      if (!fallback) {
        if (!quiet) __syncthreads();  // the previous row's windows and tables are no longer read
        quiet = false;
        // ---- column tables: meta[side][o] = LDS address of stored pair 0 of
        // the column's window (slot - lo * kB); each wave fills one side of
        // 64 consecutive columns at a time (the table writes are
        // conflict-free), so only the first nO / 32 waves spend instructions
        // on it and the pass ends after one side's work per wave.
        // Column o of side s is read at the offsets dk in [max(s, o-W+1),
        // min(o, kmax)] (one per wave), so q of every read lies between
        // sqrt(t0lo^2 + alo r^2(dk_lo)) / dt and sqrt(t0hi^2 + ahi r^2(dk_hi)) / dt
        // with the workgroup's bounds of t0^2 and a = 4/v^2.  r2tab[dk] =
        // r^2 of the row pair at crossline offset dk.
        bool ovf = false;
        // The staging jobs of a row read entries up to nO + 3G - 1 (loads
        // issued for columns past the row store into the spare slot); their
        // widths are zeroed, so those loads read nothing.
        constexpr int kPast = 4 * G + W;
        // table job tj of wave w (tj = w, w + W, ...): side tj % 2 of columns
        // 64 (tj / 2) .. 64 (tj / 2) + 63, so a row of up to 4 W - 24
        // columns takes one pass of half a wave's column work per wave
        for (int tj = w; (tj >> 1) * kWave < nO + kPast; tj += W) {
          const int side = tj & 1;
          const int o = (tj >> 1) * kWave + lane;
          char* const wid = smem + PL::kWid + (side * PL::kCols + o) * PL::kWidB;
          using I2 = int __attribute__((ext_vector_type(2)));
          if (o >= nO) {
            if (o < nO + kPast) {
              if constexpr (PL::kAdj)
                *reinterpret_cast<int*>(wid) = 0;
              else
                *reinterpret_cast<I2*>(wid) = I2{0, 0};
            }
          } else {
            int lo, wb;
            col_window(side, o, lo, wb, ovf);
            if constexpr (PL::kAdj)
              *reinterpret_cast<int*>(wid) = wb;
            else
              *reinterpret_cast<I2*>(wid) = I2{(lo * kB) >> PL::kHShift, wb};
            meta[side * PL::kCols + o] = PL::slot(side, o) - lo * kB;
          }
        }
        for (int dk = w * kWave + lane; dk <= kmax_wg; dk += W * kWave) {
          const float dyv = (float)dk * a.dy;
          r2tab[dk] = __builtin_fmaf(dyv, dyv, d2);
        }
        // (the vote runs with every lane active: lane l holds columns 64 w +
        // l, ...)
        const bool wave_ovf = __any(ovf);
        if (lane == 0) flag[w] = wave_ovf ? 1 : 0;
        if constexpr (!PL::kAdj) {
          // ---- first windows: the first G + W - 1 columns of each side, in
          // staging jobs whose windows the wave computes itself (lane t: job
          // t), so they need no barrier after the table pass; stored before
          // the one barrier that publishes tables, flags and windows (a row
          // that then falls back to the gather path leaves them unread).
          // FP32 single rows go straight into LDS, summed or scaled rows
          // through registers.
          constexpr int kFill = (2 * (G + W - 1) + W - 1) / W;
          const __amdgpu_buffer_rsrc_t rs0 = __builtin_amdgcn_make_buffer_rsrc(
              const_cast<char*>(base + (size_t)r0 * row_bytes), 0, (int)row_bytes, kRsrcWord3);
          const __amdgpu_buffer_rsrc_t rs1 = __builtin_amdgcn_make_buffer_rsrc(
              const_cast<char*>(base + (size_t)rm * row_bytes), 0, (int)row_bytes, kRsrcWord3);
          const const_f32_ptr sc0 = as_const(a.scale + (size_t)r0 * row_tr);
          const const_f32_ptr sc1 = as_const(a.scale + (size_t)rm * row_tr);
          const int nc = min(G + W - 1, nO);
          const StageJob jl = job_at(0, nc, w + min(lane, kFill - 1) * W);
          int lo_l = 0, wb_l = 0;
          bool unused_ovf = false;
          if (jl.ok) col_window(jl.side, jl.o, lo_l, wb_l, unused_ovf);
#pragma unroll
          for (int t = 0; t < kFill; ++t) {
            const StageJob jb = job_at(0, nc, w + t * W);
            if (jb.ok) {
              const int lo = __builtin_amdgcn_readlane(lo_l, t);
              const int wb = __builtin_amdgcn_readlane(wb_l, t);
              const int k = jb.side == 0 ? kbase + jb.o : kbase + (W - 1) - jb.o;
              const bool data = k >= xd_lo && k < xd_hi;
              const int soff = data ? (k - xd_lo) * tb : 0;
              const int si = data ? k - xd_lo : a.n_xd;
              const int voff = (!data || lb >= wb) ? 0x7ffffff0 : (lo * kB + lb) >> PL::kHShift;
              char* const dst = smem + PL::slot(jb.side, jb.o);
              if (S::kScaled || two) {
                F4 x = widen(load_row(rs0, voff, soff), S::kScaled ? sc0[si] : 1.0f);
                if (two) x = widen_add(load_row(rs1, voff, soff), S::kScaled ? sc1[si] : 1.0f, x);
                *reinterpret_cast<F4*>(dst + lb) = x;
              } else {
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rs0, (__attribute__((address_space(3))) void*)dst, 16, voff, soff, 0, 0);
              }
            }
          }
          __builtin_amdgcn_s_waitcnt(0);
        }
        __syncthreads();  // tables, flags and (forward) the first windows are visible
        fallback = __any(pl && flag[pl ? lane : 0] != 0);
      }
      if constexpr (PL::kAdj) {
        if (!fallback) {
          // ---- adjoint, in the forward's schedule: wave w at step dk adds the
          // terms of its lanes -- the same (i, f, w, mask) as the gather-form
          // adjoint: d[i] += w (1 - f) m, d[i + 1] += w f m -- to the windows
          // of columns w + dk (side 0) and W-1-w + dk (side 1), one tau for
          // the four pairs (+-db, +-dk).  Rows +db and -db receive the same
          // terms, so one window per column serves both.  A window holds
          // pairs (A, B): a term adds (w m, w f m) to pair i, and the flush
          // forms d[j] = A[j] - B[j] + B[j - 1].  Step u of a group adds to
          // copy u of the column's window: the W waves touch a column at W
          // consecutive steps, so within a group no two waves touch the same
          // copy, and the adds are plain read-add-write.  Column o is last
          // touched at step dk = o (wave 0 on side 0, wave W-1 on side 1), so
          // the G columns per side of a group are complete at its closing
          // barrier; during the next group their copies are summed, flushed
          // into the data cube (global atomic adds of the nonzero samples)
          // and zeroed, and their slots are reused by the columns that enter
          // the group after it (ring of 2G + W - 1 slots per side).
          const int* const wid = reinterpret_cast<const int*>(smem + PL::kWid);
          using F2 = float __attribute__((ext_vector_type(2)));
          using LdsF = __attribute__((address_space(3))) float;
          // the lane's weight factor: (1/v^2) m with amplitude correction (the
          // weight times the image sample is then rsq(tau^2) times it), m
          // without
          LaneParams pp = p;
          pp.iv2 = AMP ? p.iv2 * p.mval : p.mval;
          const bool odd = (lane & 1) != 0;
          // With one 4/v^2 over the workgroup's tile (laterally and
          // temporally constant velocity) q = sqrt(t0^2 + a r^2) / dt grows by
          // t0 / tau >= cos(theta) per lane inside the aperture, so at
          // apertures below about 60 deg (cos > 1/2 with a margin) no two
          // valid lanes of equal parity share an i: no check.
          const bool mono = gAlo == gAhi && a.alpha_mask * a.alpha_mask * a.tan2 <
                                                2.9f * (a.alpha * a.alpha);
          // pair i of both columns (one: side 0 only; the two never coincide)
          auto rmw = [&](int ar, int al, bool one, float wm, float wf) {
            F2* const cr = reinterpret_cast<F2*>(smem + ar);
            F2* const cl = reinterpret_cast<F2*>(smem + al);
            const F2 vr = *cr;
            const F2 vl = one ? F2{0.0f, 0.0f} : *cl;
            *cr = F2{vr.x + wm, vr.y + wf};
            if (!one) *cl = F2{vl.x + wm, vl.y + wf};
          };
          auto lds_add = [&](int addr, float v) {
            __hip_atomic_fetch_add((LdsF*)(smem + addr), v, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
          };
          // flush column o of `side` into each row: lane l forms the column's
          // samples lo + l and lo + l + 64 from the G copies (each atomic
          // instruction covers 256 contiguous bytes), then zeroes the copies
          auto flush = [&](int side, int o) {
            const int sl = PL::slot(side, o);
            const int lo = uniform_i((sl - meta[side * PL::kCols + o]) / 8);
            const int n = uniform_i(wid[side * PL::kCols + o]);
            const int k = side == 0 ? kbase + o : kbase + (W - 1) - o;
            float v0 = 0.0f, v1 = 0.0f;
#pragma unroll
            for (int c = 0; c < G; ++c) {
              const F2* const win = reinterpret_cast<const F2*>(smem + sl + c * PL::kSlot);
              const F2 p0 = win[lane], p1 = win[lane + kWave];
              const float bm0 = lane > 0 ? win[lane - 1].y : 0.0f;
              const float bm1 = win[lane + kWave - 1].y;
              v0 += (p0.x - p0.y) + bm0;
              v1 += (p1.x - p1.y) + bm1;
            }
            asm volatile("" ::: "memory");  // every read before the zeroing
#pragma unroll
            for (int c = 0; c < G; ++c) {
              F2* const win = reinterpret_cast<F2*>(smem + sl + c * PL::kSlot);
              win[lane] = F2{0.0f, 0.0f};
              win[lane + kWave] = F2{0.0f, 0.0f};
            }
            if (k >= xd_lo && k < xd_hi && a.adj_rows) {
              // the window once, into the pass's row-pair buffer (added into
              // both rows of the pair by launch_adjoint_rows)
              float* const dst =
                  a.adj_rows +
                  (((size_t)io * a.adj_nd + (d - a.d_lo)) * a.n_xd + (k - xd_lo)) * a.n_t + lo +
                  lane;
              if (lane < n && v0 != 0.0f) unsafeAtomicAdd(dst, v0);
              if (lane + kWave < n && v1 != 0.0f) unsafeAtomicAdd(dst + kWave, v1);
            } else if (k >= xd_lo && k < xd_hi) {  // outside the data: nothing scattered there counts
              float* const dst0 = a.adj_out + ((size_t)r0 * a.n_xd + (k - xd_lo)) * a.n_t + lo + lane;
              if (lane < n && v0 != 0.0f) unsafeAtomicAdd(dst0, v0);
              if (lane + kWave < n && v1 != 0.0f) unsafeAtomicAdd(dst0 + kWave, v1);
              if (two) {
                float* const dst1 = a.adj_out + ((size_t)rm * a.n_xd + (k - xd_lo)) * a.n_t + lo + lane;
                if (lane < n && v0 != 0.0f) unsafeAtomicAdd(dst1, v0);
                if (lane + kWave < n && v1 != 0.0f) unsafeAtomicAdd(dst1 + kWave, v1);
              }
            }
          };
          // flush the columns [oa, ob) of both sides (jobs spread over the waves)
          auto flush_cols = [&](int oa, int ob) {
            for (int jb = w; jb < 2 * (ob - oa); jb += W) flush(jb & 1, oa + (jb >> 1));
          };
          using Tab = StageTab<G, 1>;
          auto read_tab = [&](int dk0, Tab& t) {
#pragma unroll
            for (int u = 0; u < G; ++u) {
              const int dk = dk0 + u;
              t.r2[u] = r2tab[dk];
              t.bR[u] = meta[w + dk] + u * PL::kSlot;
              t.bL[u] = meta[PL::kCols + (W - 1 - w) + dk] + u * PL::kSlot;
            }
          };
          // steps dk0 + u, u < n (FULL: all G); FIRST: dk = 0 has one column
          auto compute = [&](auto full_c, auto first_c, auto mask_c, int dk0, int n, const Tab& t) {
            constexpr bool MASK = decltype(mask_c)::value;
            constexpr bool FULL = decltype(full_c)::value;
            constexpr bool FIRST = decltype(first_c)::value;
#pragma unroll
            for (int u = 0; u < G; ++u) {
              if (FULL || u < n) {
                const bool one = FIRST && u == 0;
                const float r2 = t.r2[u];
                const TauW tw = tau_weight(pp, r2, AMP);
                const float q = tw.tau * a.inv_dt;
                const int iq = (int)q;
                const int i = min(iq, nt1);  // (inside the window; terms at n_t - 1 are off)
                bool on = valid && iq < nt1;
                if constexpr (MASK) on = on && r2 <= p.R2;
                const float f = __builtin_amdgcn_fractf(q);
                const float wm = AMP ? tw.w : pp.iv2;
                const float wf = wm * f;
                int ar, al;
                asm("v_lshl_add_u32 %0, %1, 3, %2" : "=v"(ar) : "v"(i), "v"(t.bR[u]));
                asm("v_lshl_add_u32 %0, %1, 3, %2" : "=v"(al) : "v"(i), "v"(t.bL[u]));
                bool fast = mono;
                if (!mono) {
                  // i nondecreasing over the valid lanes with at most two
                  // lanes per value (lanes without a sample sort last)
                  const int ik = valid ? iq : 0x40000000 + lane;
                  const int ip = lane_prev(ik), ipp = lane_prev(ip);
                  fast = !__any((lane > 0 && ik < ip) || (lane > 1 && ik == ip && ip == ipp));
                }
                if (fast) {
                  // two passes, even lanes then odd lanes: within a pass no
                  // two lanes share a pair
                  if (on && !odd) rmw(ar, al, one, wm, wf);
                  asm volatile("" ::: "memory");
                  if (on && odd) rmw(ar, al, one, wm, wf);
                } else if (on) {
                  // lanes sharing a pair: LDS atomics (the copy is this wave's alone)
                  lds_add(ar, wm);
                  lds_add(ar + 4, wf);
                  if (!one) {
                    lds_add(al, wm);
                    lds_add(al + 4, wf);
                  }
                }
                asm volatile("" ::: "memory");
              }
            }
          };
          // one group: flush the previous group's columns, add this group's
          // terms, read the next group's tables; the barrier completes the
          // group's adds and the flushes' zeroing
          auto run_group = [&](auto full_c, auto first_c, auto mask_c, int dk0, Tab& t) {
            if (dk0 > 0) flush_cols(dk0 - G, min(dk0, nO));
            compute(full_c, first_c, mask_c, dk0, min(G, kmax_wg + 1 - dk0), t);
            read_tab(dk0 + G, t);
            __syncthreads();
          };
          using T = std::integral_constant<bool, true>;
          using F = std::integral_constant<bool, false>;
          Tab t;
          read_tab(0, t);
          if (kmax_wg + 1 >= G)
            run_group(T{}, T{}, T{}, 0, t);
          else
            run_group(F{}, T{}, T{}, 0, t);
          int dk0 = G;
          // groups inside every valid lane's aperture (dk <= kin) skip the mask
          for (; dk0 + G <= min(kmax_wg, kin) + 1; dk0 += G) run_group(T{}, F{}, F{}, dk0, t);
          for (; dk0 + G <= kmax_wg + 1; dk0 += G) run_group(T{}, F{}, T{}, dk0, t);
          if (dk0 <= kmax_wg) {
            run_group(F{}, F{}, T{}, dk0, t);
            dk0 += G;
          }
          // the last group's columns and every column after it
          flush_cols(dk0 - G, nO);
          continue;
        }
      } else if (!fallback) {
        // ---- staging jobs: job j of a set of columns [oa, oa + nc) is
        // (column oa + j / 2, side j % 2); wave w takes jobs w, w + W, ...
        // A job loads the column's window of both rows of the pair, two pairs
        // per lane (16 bytes of FP32 pairs, 8 of half pairs, per row), and
        // stores their sum: in the first fill of a row at once, later through
        // registers that are written to the column's ring slot at the end of
        // a group, so LDS reads never wait on loads in flight.  Pairs past
        // the trace end are loaded out of range; columns outside the data
        // are zeros (out of range in the first fill, the row's zero trace
        // afterwards).
        const int lbh = lb >> PL::kHShift;  // the lane's cube byte offset in a window
        const __amdgpu_buffer_rsrc_t rs0 = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<char*>(base + (size_t)r0 * row_bytes), 0, (int)row_bytes, kRsrcWord3);
        const __amdgpu_buffer_rsrc_t rs1 = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<char*>(base + (size_t)rm * row_bytes), 0, (int)row_bytes, kRsrcWord3);
        // scale tables of the staged rows (F16 / BF16; entry n_xd is the
        // row's zero trace)
        const const_f32_ptr sc0 = as_const(a.scale + (size_t)r0 * row_tr);
        const const_f32_ptr sc1 = as_const(a.scale + (size_t)rm * row_tr);
        auto tab = [&](int side, int o) { return meta[side * PL::kCols + o]; };
        // a column's window entry: (lo * kHB, bytes of pairs lo .. hi in LDS)
        auto wtab = [&](int side, int o) {
          using I2 = int __attribute__((ext_vector_type(2)));
          return *reinterpret_cast<const I2*>(smem + PL::kWid + (side * PL::kCols + o) * PL::kWidB);
        };
        // Table values of one group, read one group ahead so that their LDS
        // latency overlaps the stores and the barrier: r^2, the two column
        // addresses of each step, and the table entries of the staging jobs
        // for the group after it.
        // staging jobs per wave per group, at most
        constexpr int kJobs = (2 * G + W - 1) / W;
        // This wave's jobs in every full group of the row: job q is column
        // oa + jc of side js if the wave has that job at all (oa = the
        // group's first new column), so the jobs' state advances by fixed
        // steps from group to group and is kept in scalar registers: jo =
        // column offset of the next load, jsl = its ring slot, jkd = its data
        // crossline (buffer crossline - xd_lo), jso = jkd * tb.  A column
        // outside the data loads the row's zero trace (its last); a job past
        // the row's columns (or a wave without the job) loads a column like
        // any other and stores into the spare slot.
        int js[kJobs], jo[kJobs], jsl[kJobs], jkd[kJobs], jso[kJobs];
        int jlim[kJobs], jslend[kJobs], jdkd[kJobs];
#pragma unroll
        for (int q = 0; q < kJobs; ++q) {
          const StageJob jb = job_at(0, G, w + q * W);
          const int jc = uniform_i(jb.o);
          js[q] = uniform_i(jb.side);
          jo[q] = G + W - 1 + jc;
          jlim[q] = jb.ok ? nO : -1;
          jsl[q] = PL::slot(js[q], jo[q]);
          jslend[q] = js[q] * PL::kSide + PL::kRing * PL::kSlot;
          jkd[q] = (js[q] == 0 ? kbase + jo[q] : kbase + (W - 1) - jo[q]) - xd_lo;
          jdkd[q] = js[q] == 0 ? G : -G;
          jso[q] = jkd[q] * tb;
        }
        using Tab = StageTab<G, kJobs>;
        auto read_tab = [&](int dk0, Tab& t) {
#pragma unroll
          for (int u = 0; u < G; ++u) {
            const int dk = dk0 + u;
            t.r2[u] = r2tab[dk];
            t.bR[u] = tab(0, w + dk);
            t.bL[u] = tab(1, (W - 1 - w) + dk);
          }
          // table entries of the next loads (read whether or not the job
          // exists: the index stays inside the table)
#pragma unroll
          for (int q = 0; q < kJobs; ++q) {
            const auto e = wtab(js[q], jo[q]);
            t.mj[q] = e.x;
            t.mw[q] = e.y;
          }
        };
        // Steps dk0 + u, u < n (FULL: all G).  Every wave sweeps the
        // workgroup's offsets: each column's window covers all waves' reads up
        // to kmax_wg, and a step past a wave's own last offset adds only masked
        // terms or exact zeros (pairs past the trace end).  FIRST: the group
        // holds dk = 0, whose -dk column is the +dk column and is left out.
        auto compute = [&](auto full_c, auto first_c, auto mask_c, int dk0, int n, const Tab& t) {
          using LS = Store<PSM_STORE_F32>;  // the windows hold FP32 pairs
          constexpr bool MASK = decltype(mask_c)::value;  // false: every lane is inside
          constexpr bool FULL = decltype(full_c)::value;
          constexpr bool FIRST = decltype(first_c)::value;
          struct Stp {
            float f, wt;
            bool m;
            LS::Raw x[2];
          };
          Stp st[G];
#pragma unroll
          for (int u = 0; u < G; ++u) {
            if (FULL || u < n) {
              const bool one = FIRST && u == 0;  // dk = 0: one column
              const float r2 = t.r2[u];
              const TauW tw = tau_weight(p, r2, AMP);
              const float q = tw.tau * a.inv_dt;
              // both column addresses as (i << log2 kB) + base, one
              // instruction each (the compiler shares the shift otherwise)
              // (a lane past the trace end reads a zero pair of the window)
              const int i = (int)q;
              int ar, al;
              asm("v_lshl_add_u32 %0, %1, %2, %3" : "=v"(ar) : "v"(i), "i"(LS::kShift), "v"(t.bR[u]));
              asm("v_lshl_add_u32 %0, %1, %2, %3" : "=v"(al) : "v"(i), "i"(LS::kShift), "v"(t.bL[u]));
              st[u].x[0] = LS::load_ptr(smem + ar);
              st[u].x[1] = one ? LS::Raw{} : LS::load_ptr(smem + al);
              st[u].f = __builtin_amdgcn_fractf(q);
              st[u].wt = tw.w;
              if constexpr (MASK) st[u].m = r2 <= p.R2;
            }
          }
          // each window holds the sum of the rows +db and -db: columns +dk
          // and -dk
#pragma unroll
          for (int u = 0; u < G; ++u) {
            if (FULL || u < n) {
              const Stp& q = st[u];
              const bool one = FIRST && u == 0;
              float sum = LS::lerp(q.x[0], q.f);
              if (!one) sum = sum + LS::lerp(q.x[1], q.f);
              if constexpr (MASK) sum = q.m ? sum : 0.0f;
              acc = __builtin_fmaf(sum, q.wt, acc);
            }
          }
        };
        // one group: the next group's G new columns are loaded while this
        // group runs and stored into slots this group does not read (double
        // ring) or, after a barrier, into the slots of the G columns no later
        // group reads (single ring).  Every wave issues the same
        // loads whether or not its jobs exist (absent jobs read out of range),
        // so no loaded value is merged across a branch and nothing waits on
        // the loads before the stores.
        // The G columns a group adds to the ring are loaded two groups ahead
        // of their reads: the loads issued at the start of group dk0 are held
        // in registers through groups dk0 and dk0 + G, stored at the end of
        // group dk0 + G and read from group dk0 + 2G on.
        // A register set: the loads of this wave's jobs (v[q] the first row,
        // v[kJobs + q] the second row of a pair), their destination slots and
        // their scales (F16 / BF16).
        using Jobs = JV[2 * kJobs];
        using Dsts = int[kJobs];
        using Scls = float[2 * kJobs];
        auto issue = [&](auto nr_c, const int* m, const int* mw, Jobs& v, Dsts& dst, Scls& sc) {
          constexpr int NR = decltype(nr_c)::value;
#pragma unroll
          for (int q = 0; q < kJobs; ++q) {
            // a job loads pairs lo .. lo + kWinPairs - 1 of its column (m =
            // lo * kHB, the window's cube byte offset); lanes whose pairs all lie past the
            // trace end load out of range (no read reaches those slot
            // entries: reads stop at hi <= n_t - 1), the lane holding pair
            // n_t - 1 ends its load there (lo makes n_t - 1 - lo odd)
            const bool data = (unsigned)jkd[q] < (unsigned)a.n_xd;
            dst[q] = jo[q] < jlim[q] ? jsl[q] : PL::kSpare;
            // (lanes past the window's pairs lo .. hi -- hi <= n_t - 1 --
            // load nothing)
            const int vo = lb < mw[q] ? m[q] + lbh : 0x7ff00000;
            const int so = data ? jso[q] : a.n_xd * tb;
            v[q] = load_row(rs0, vo, so);
            if constexpr (NR == 2) v[kJobs + q] = load_row(rs1, vo, so);
            if constexpr (S::kScaled) {
              const int si = data ? jkd[q] : a.n_xd;
              sc[q] = sc0[si];
              if constexpr (NR == 2) sc[kJobs + q] = sc1[si];
            }
            jo[q] += G;
            jkd[q] += jdkd[q];
            jso[q] += jdkd[q] * tb;
            jsl[q] += G * PL::kSlot;
            if (jsl[q] >= jslend[q]) jsl[q] -= PL::kRing * PL::kSlot;
          }
        };
        // one group: issue the loads for group dk0 + 2G (into vi), sum the
        // group, read the next group's tables, store the loads issued one
        // group earlier (vs) and make them visible.
        auto run_group = [&](auto nr_c, auto full_c, auto first_c, auto mask_c, int dk0, Tab& t,
                             Jobs& vs, Dsts& ds, Scls& ss, Jobs& vi, Dsts& di, Scls& si) {
          // (wave priority: the loads go out first and the group's tail --
          // stores and barrier -- is not held up behind other waves'
          // sums; 0.75 % on the production survey)
          __builtin_amdgcn_s_setprio(1);
          issue(nr_c, t.mj, t.mw, vi, di, si);
          __builtin_amdgcn_s_setprio(0);
          constexpr int NR = decltype(nr_c)::value;
          compute(full_c, first_c, mask_c, dk0, min(G, kmax_wg + 1 - dk0), t);
          read_tab(dk0 + G, t);
          // The stores wait on the loads, so the group's sum must be complete
          // before them: the empty asm takes the accumulator as an input, and
          // no store is conditional (absent jobs write the spare slot), so no
          // summation is sunk past a store.
          asm volatile("" ::"v"(acc));
          __builtin_amdgcn_s_setprio(2);
          __builtin_amdgcn_sched_barrier(0);
          if constexpr (PL::kSingle) __syncthreads();  // this group's reads are done
#pragma unroll
          for (int q = 0; q < kJobs; ++q) {
            if (lb < PL::kSlot) {
              F4 x = widen(vs[q], S::kScaled ? ss[q] : 1.0f);
              if constexpr (NR == 2)
                x = widen_add(vs[kJobs + q], S::kScaled ? ss[kJobs + q] : 1.0f, x);
              *reinterpret_cast<F4*>(smem + ds[q] + lb) = x;
            }
          }
          __syncthreads();  // the next group's windows are visible
          __builtin_amdgcn_s_setprio(0);
        };
        auto run_row = [&](auto nr_c) {
          using T = std::integral_constant<bool, true>;
          using F = std::integral_constant<bool, false>;
          Jobs va, vb;
          Dsts da, db;
          Scls sa, sb;
          {  // the loads stored at the end of group 0
            int m0[kJobs], w0[kJobs];
#pragma unroll
            for (int q = 0; q < kJobs; ++q) {
              const auto e = wtab(js[q], jo[q]);
              m0[q] = e.x;
              w0[q] = e.y;
            }
            issue(nr_c, m0, w0, va, da, sa);
          }
          Tab t;
          read_tab(0, t);
          // (the first windows were published by the barrier after the
          // table pass)
          if (kmax_wg + 1 >= G)
            run_group(nr_c, T{}, T{}, T{}, 0, t, va, da, sa, vb, db, sb);
          else
            run_group(nr_c, F{}, T{}, T{}, 0, t, va, da, sa, vb, db, sb);
          // groups alternate the two register sets (unrolled by two); pairs
          // of groups inside every lane's aperture (dk <= kin) skip the mask
          // (the barriers are the same on both paths)
          int dk0 = G;
          for (; dk0 + 2 * G <= min(kmax_wg, kin) + 1; dk0 += 2 * G) {
            run_group(nr_c, T{}, F{}, F{}, dk0, t, vb, db, sb, va, da, sa);
            run_group(nr_c, T{}, F{}, F{}, dk0 + G, t, va, da, sa, vb, db, sb);
          }
          for (; dk0 + 2 * G <= kmax_wg + 1; dk0 += 2 * G) {
            run_group(nr_c, T{}, F{}, T{}, dk0, t, vb, db, sb, va, da, sa);
            run_group(nr_c, T{}, F{}, T{}, dk0 + G, t, va, da, sa, vb, db, sb);
          }
          if (dk0 + G <= kmax_wg + 1) {
            run_group(nr_c, T{}, F{}, T{}, dk0, t, vb, db, sb, va, da, sa);
            if (dk0 + G <= kmax_wg) run_group(nr_c, F{}, F{}, T{}, dk0 + G, t, va, da, sa, vb, db, sb);
          } else if (dk0 <= kmax_wg) {
            run_group(nr_c, F{}, F{}, T{}, dk0, t, vb, db, sb, va, da, sa);
          }
        };
        if (two)
          run_row(std::integral_constant<int, 2>{});
        else
          run_row(std::integral_constant<int, 1>{});
        quiet = true;
        continue;
      }
      // gather path for this row
      if (R2w >= d2)
        sweep_row<ST, false, MODE>(a, kj, p, R2w, acc, cap, cg,
                                       hp ? base + (size_t)rp * row_bytes : zero_row, hp,
                                       hp ? rp : a.n_rows,
                                       hm ? base + (size_t)rm * row_bytes : zero_row, hm,
                                       hm ? rm : a.n_rows, d2);
    }
    if constexpr (MODE == kForward && AMP) acc *= lane_iv2;
  }
  if constexpr (MODE == kForward) {
    // the lane's output sample, formed again from the lane index (which the
    // aperture loop keeps live) rather than held through the loop
    int le = lane;
    asm volatile("" : "+v"(le));
    const int so = tile * kWave + le;
    if (in_out && so < a.n_t) {
      float* orow = a.out + ((size_t)io * a.n_out_xl + j) * a.n_t;
      if (!in_buf) {
        if (!a.accumulate) orow[so] = 0.0f;  // migration_kernel.cpp:79-85
      } else if (a.accumulate) {
        if (valid) orow[so] += acc;
      } else {
        orow[so] = valid ? acc : 0.0f;  // :148-151, :375
      }
    }
  }
}

// The staged form over a grid of (t0 tile, output inline, crossline group)
// workgroups: latest tiles first, consecutive spatial indices of a tile on one
// XCD (as migrate_kernel).
template <int ST, int W, int G, bool AMP>
__global__ __launch_bounds__(kWave * W)
__attribute__((amdgpu_waves_per_eu(LdsPlan<ST, W, G>::kWavesPerEu)))
void migrate_lds_kernel(MigArgs a, int n_tiles, int n_jgroups, int per_tile_pad) {
  __shared__ __attribute__((aligned(16))) char smem[LdsPlan<ST, W, G>::kBytes];
  const int g = blockIdx.x;
  const int tile = n_tiles - 1 - g / per_tile_pad;
  const int rem = g % per_tile_pad;
  const int spatial = (rem % kXcds) * (per_tile_pad / kXcds) + rem / kXcds;
  if (spatial >= a.n_out_il * n_jgroups) return;  // whole workgroup
  const int io = spatial / n_jgroups;
  staged_tile<ST, W, G, AMP>(a, a.tasks[io], tile, io, (spatial % n_jgroups) * W, smem);
}

// The staged adjoint over the same grid as migrate_lds_kernel (FP32 data).
template <int W, int G, bool AMP>
__global__ __launch_bounds__(kWave * W)
void demigrate_lds_kernel(MigArgs a, int n_tiles, int n_jgroups, int per_tile_pad) {
  __shared__ __attribute__((aligned(16))) char smem[LdsPlan<PSM_STORE_F32, W, G, kAdjoint>::kBytes];
  const int g = blockIdx.x;
  const int tile = n_tiles - 1 - g / per_tile_pad;
  const int rem = g % per_tile_pad;
  const int spatial = (rem % kXcds) * (per_tile_pad / kXcds) + rem / kXcds;
  if (spatial >= a.n_out_il * n_jgroups) return;  // whole workgroup
  const int io = spatial / n_jgroups;
  staged_tile<PSM_STORE_F32, W, G, AMP, kAdjoint>(a, a.tasks[io], tile, io,
                                                 (spatial % n_jgroups) * W, smem);
}

// One dword of a read-only table through the constant address space (a
// scalar load at a wave-uniform index).
template <class T>
__device__ __forceinline__ int const_dword(const T* base, int idx) {
  using CI = const __attribute__((address_space(4))) int*;
  return ((CI)base)[idx];
}

// The next work item of a persistent batch workgroup on XCD `xcd`: from that
// XCD's queue first, then from the other XCDs' queues in turn (`pass` counts
// the queues found drained).  Returns false once every queue is drained.  All
// waves of the workgroup call it together; s_next is an LDS word.  The queue
// position comes from LDS and is made uniform explicitly, so everything
// derived from it -- the segment, the item's arguments and buffer resources
// -- is read through scalar loads and stays in scalar registers.
__device__ __forceinline__ bool batch_next(const BatchSeg* segs, int n_segs, unsigned* counters,
                                           const unsigned* totals, int xcd, int& pass,
                                           unsigned* s_next, BatchSeg& sg, int& spatial) {
  constexpr int kXfirst = (int)(offsetof(BatchSeg, xfirst) / 4);
  for (; pass < kBatchXcds; ++pass) {
    const int x = uniform_i((xcd + pass) % kBatchXcds);
    const unsigned total = (unsigned)const_dword(totals, x);
    if (threadIdx.x == 0) *s_next = total ? atomicAdd(&counters[x], 1u) : 0u;
    __syncthreads();
    const unsigned wi = (unsigned)uniform_i((int)*s_next);
    __syncthreads();
    if (wi >= total) continue;
    int lo = 0, hi = n_segs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      const int f = const_dword(segs + mid, kXfirst + x);
      if ((unsigned)f <= wi) lo = mid; else hi = mid - 1;
    }
    sg = const_load<BatchSeg>(segs + lo);
    spatial = const_dword(segs + lo, (int)(offsetof(BatchSeg, xbase) / 4) + x) +
              (int)(wi - (unsigned)const_dword(segs + lo, kXfirst + x));
    return true;
  }
  return false;
}

// Persistent batch kernel: every workgroup takes work items (one
// (volume, t0 tile, crossline group) each) from the queues until all are
// drained, so volumes of very different size and cost share one launch.
// Each queue is ordered by segment: segment k = the groups of one t0 tile of
// one volume, latest tiles of all volumes first.
template <int ST, int WPB>
__global__ __launch_bounds__(kWave * WPB) void batch_kernel(const MigArgs* __restrict__ items,
                                                            const BatchSeg* __restrict__ segs,
                                                            int n_segs,
                                                            unsigned* __restrict__ counters,
                                                            const unsigned* __restrict__ totals) {
  __shared__ unsigned s_next;
  const int xcd = (int)(blockIdx.x % kBatchXcds);
  int pass = 0;
  BatchSeg sg;
  int spatial;
  while (batch_next(segs, n_segs, counters, totals, xcd, pass, &s_next, sg, spatial)) {
    const MigArgs a = const_load<MigArgs>(items + sg.item);
    const int io = spatial / sg.n_jgroups;
    const int j = (spatial % sg.n_jgroups) * WPB + uniform_i((int)(threadIdx.x / kWave));
    if (io < a.n_out_il && j < a.n_out_xl) migrate_trace_tile<ST, false, kForward>(a, sg.tile, io, j);
  }
}

// The same queues over workgroups of the staged form (W output crosslines per
// item): each item runs staged_tile, so its result equals migrate_lds_kernel's
// for that (tile, inline, crossline group) bit for bit.  The barriers around
// the queue read also separate one item's LDS use from the next.
// (registers capped for the residency the LDS allows: the item's scalar
// state otherwise spills into VGPR lanes past that)
template <int ST, int W, int G>
__global__ __launch_bounds__(kWave * W)
__attribute__((amdgpu_waves_per_eu(LdsPlan<ST, W, G>::kWavesPerEu)))
void batch_lds_kernel(const MigArgs* __restrict__ items, const BatchSeg* __restrict__ segs,
                      int n_segs, unsigned* __restrict__ counters,
                      const unsigned* __restrict__ totals) {
  __shared__ __attribute__((aligned(16))) char smem[LdsPlan<ST, W, G>::kBytes];
  __shared__ unsigned s_next;
  const int xcd = (int)(blockIdx.x % kBatchXcds);
  int pass = 0;
  BatchSeg sg;
  int spatial;
  while (batch_next(segs, n_segs, counters, totals, xcd, pass, &s_next, sg, spatial)) {
    const MigArgs a = const_load<MigArgs>(items + sg.item);
    const int io = spatial / sg.n_jgroups;
    const int j0 = (spatial % sg.n_jgroups) * W;
    if (io < a.n_out_il) {
      // (the task through scalar loads too: the row indices derive from it)
      const Task tk = const_load<Task>(a.tasks + io);
      if (a.amp)
        staged_tile<ST, W, G, true>(a, tk, sg.tile, io, j0, smem);
      else
        staged_tile<ST, W, G, false>(a, tk, sg.tile, io, j0, smem);
    }
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

// Row-pair windows of one staged-adjoint pass into the data cube: the block
// (r, k, 256-sample chunk) sums the planes plane[e] of the row-pair buffer
// listed for data row r (entries start[r] .. start[r + 1] - 1, each an output
// inline and row offset whose pair holds row r, in output-inline order) and
// adds the sum into data[r][k][chunk].  Each sample has one thread (no
// atomics); the list is uniform and the loads independent.
__global__ __launch_bounds__(256) void adjoint_rows_kernel(
    const float* __restrict__ rows, float* __restrict__ data, const int* __restrict__ start,
    const int* __restrict__ plane, int n_rows, int n_xd, int n_t, int n_chunks) {
  const int blk = blockIdx.x;
  const int chunk = blk % n_chunks;
  const int rk = blk / n_chunks;
  const int r = rk / n_xd, k = rk % n_xd;
  const int i = chunk * 256 + (int)threadIdx.x;
  if (r >= n_rows || i >= n_t) return;
  const int e0 = start[r], e1 = start[r + 1];
  const size_t pstride = (size_t)n_xd * n_t;
  const float* const src = rows + (size_t)k * n_t + i;
  float acc = 0.0f;
#pragma unroll 8
  for (int e = e0; e < e1; ++e) acc += src[(size_t)plane[e] * pstride];
  float* const o = data + ((size_t)r * n_xd + k) * n_t + i;
  *o += acc;
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

template <int ST, int W, int G>
hipError_t launch_lds_wg(const MigArgs& a, hipStream_t stream) {
  const int n_tiles = (a.n_t + kWave - 1) / kWave;
  const int n_jgroups = (a.n_out_xl + W - 1) / W;
  const long long per_tile = (long long)a.n_out_il * n_jgroups;
  const long long per_tile_pad = (per_tile + kXcds - 1) / kXcds * kXcds;
  const long long blocks = per_tile_pad * n_tiles;
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  if (a.amp)
    hipLaunchKernelGGL((migrate_lds_kernel<ST, W, G, true>), dim3((unsigned)blocks),
                       dim3(kWave * W), 0, stream, a, n_tiles, n_jgroups, (int)per_tile_pad);
  else
    hipLaunchKernelGGL((migrate_lds_kernel<ST, W, G, false>), dim3((unsigned)blocks),
                       dim3(kWave * W), 0, stream, a, n_tiles, n_jgroups, (int)per_tile_pad);
  return hipGetLastError();
}

// staged adjoint: kAdjWaves output crosslines per workgroup (each column
// window collects the terms of all of them before its one flush), groups of
// kAdjSteps crossline offsets
constexpr int kAdjWaves = 8;
constexpr int kAdjSteps = 2;
hipError_t launch_adj_lds(const MigArgs& a, hipStream_t stream) {
  constexpr int W = kAdjWaves, G = kAdjSteps;
  const int n_tiles = (a.n_t + kWave - 1) / kWave;
  const int n_jgroups = (a.n_out_xl + W - 1) / W;
  const long long per_tile = (long long)a.n_out_il * n_jgroups;
  const long long per_tile_pad = (per_tile + kXcds - 1) / kXcds * kXcds;
  const long long blocks = per_tile_pad * n_tiles;
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  if (a.amp)
    hipLaunchKernelGGL((demigrate_lds_kernel<W, G, true>), dim3((unsigned)blocks), dim3(kWave * W),
                       0, stream, a, n_tiles, n_jgroups, (int)per_tile_pad);
  else
    hipLaunchKernelGGL((demigrate_lds_kernel<W, G, false>), dim3((unsigned)blocks),
                       dim3(kWave * W), 0, stream, a, n_tiles, n_jgroups, (int)per_tile_pad);
  return hipGetLastError();
}

bool adjoint_staged(const MigArgs& a, const Tuning& tune) {
  return a.row_pos == nullptr && tune.staging != 0 &&
         (size_t)(a.n_xd + 1) * a.n_t * 4 <= kLdsMaxRowBytes;
}

hipError_t launch_adjoint_staged(const MigArgs& a, hipStream_t stream) {
  if (a.n_out_il <= 0 || a.n_out_xl <= 0 || a.n_t <= 0) return hipSuccess;
  return launch_adj_lds(a, stream);
}

hipError_t launch_adjoint_rows(const float* rows, float* data, const int* start, const int* plane,
                               int n_rows, int n_xd, int n_t, hipStream_t stream) {
  if (n_rows <= 0 || n_xd <= 0 || n_t <= 0) return hipSuccess;
  const int n_chunks = (n_t + 255) / 256;
  const long long blocks = (long long)n_rows * n_xd * n_chunks;
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(adjoint_rows_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, rows,
                     data, start, plane, n_rows, n_xd, n_t, n_chunks);
  return hipGetLastError();
}

template <int ST>
hipError_t launch_lds_st(const MigArgs& a, hipStream_t stream, int wpb, int steps) {
  if (wpb == 8)
    return steps == 2 ? launch_lds_wg<ST, 8, 2>(a, stream) : launch_lds_wg<ST, 8, 4>(a, stream);
  return steps == 2 ? launch_lds_wg<ST, 4, 2>(a, stream) : launch_lds_wg<ST, 4, 4>(a, stream);
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
  // LDS-staged forward kernel: contiguous rows, 4 or 8 output crosslines per
  // workgroup, rows addressable by a 32-bit buffer size
  const size_t row_bytes = (size_t)(a0.n_xd + 1) * a0.n_t * (a0.storage == PSM_STORE_F32 ? 8 : 4);
  const bool lds_ok = mode == kForward && a0.row_pos == nullptr && row_bytes <= kLdsMaxRowBytes;
  const bool auto_lds = tune.staging < 0 && lds_ok;
  const int wpb = auto_lds ? kAutoStageWaves : tune.waves_per_block;
  const int steps = auto_lds ? kAutoStageSteps : tune.stage_steps;
  const bool lds = auto_lds || (tune.staging > 0 && lds_ok && (wpb == 4 || wpb == 8));
  // staged adjoint: contiguous rows, unless staging is switched off
  const bool adj_lds = mode == kAdjoint && adjoint_staged(a0, tune);
  // automatic passes: 8 row offsets per launch for FP32 gathers, one launch otherwise
  const int rpp = tune.rows_per_pass >= 0 ? tune.rows_per_pass
                                          : ((f32_traces && !lds && !adj_lds) ? kAutoRowsF32 : 0);
  const int R = (mode == kCount || rpp <= 0) ? (row_extent + 1) : rpp;
  for (int lo = 0, pass = 0; lo <= row_extent; lo += R, ++pass) {
    a.d_lo = lo;
    a.d_hi = lo + R;
    a.accumulate = pass > 0 ? 1 : 0;
    hipError_t e;
    switch (mode) {
      case kCount: e = launch_st<PSM_STORE_F32, kCount>(a, stream, wpb); break;
      case kAdjoint:
        e = adj_lds ? launch_adj_lds(a, stream) : launch_st<PSM_STORE_F32, kAdjoint>(a, stream, wpb);
        break;
      case kTangent: e = launch_st<PSM_STORE_F32, kTangent>(a, stream, wpb); break;
      default:
        if (lds) {
          switch (a.storage) {
            case PSM_STORE_F16: e = launch_lds_st<PSM_STORE_F16>(a, stream, wpb, steps); break;
            case PSM_STORE_BF16: e = launch_lds_st<PSM_STORE_BF16>(a, stream, wpb, steps); break;
            default: e = launch_lds_st<PSM_STORE_F32>(a, stream, wpb, steps); break;
          }
          break;
        }
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

template <int ST, int W, int G>
void launch_batch_lds(const MigArgs* items, const BatchSeg* segs, int n_segs, unsigned* counters,
                      const unsigned* totals, int blocks, hipStream_t stream) {
  hipLaunchKernelGGL((batch_lds_kernel<ST, W, G>), dim3((unsigned)blocks), dim3(kWave * W), 0,
                     stream, items, segs, n_segs, counters, totals);
}

template <int ST>
hipError_t launch_batch_st(bool staged, int stage_g, const MigArgs* items, const BatchSeg* segs,
                           int n_segs, unsigned* counters, const unsigned* totals, int wpb,
                           int blocks, hipStream_t stream) {
  const dim3 grid((unsigned)blocks);
  if (staged) {
    // the shapes the batch autotuner chooses from
    if (wpb == 4)
      (stage_g == 2 ? launch_batch_lds<ST, 4, 2> : launch_batch_lds<ST, 4, 4>)(
          items, segs, n_segs, counters, totals, blocks, stream);
    else
      (stage_g == 2 ? launch_batch_lds<ST, 8, 2> : launch_batch_lds<ST, 8, 4>)(
          items, segs, n_segs, counters, totals, blocks, stream);
    return hipGetLastError();
  }
  switch (wpb) {
    case 1: hipLaunchKernelGGL((batch_kernel<ST, 1>), grid, dim3(kWave), 0, stream, items, segs, n_segs, counters, totals); break;
    case 2: hipLaunchKernelGGL((batch_kernel<ST, 2>), grid, dim3(2 * kWave), 0, stream, items, segs, n_segs, counters, totals); break;
    case 8: hipLaunchKernelGGL((batch_kernel<ST, 8>), grid, dim3(8 * kWave), 0, stream, items, segs, n_segs, counters, totals); break;
    default: hipLaunchKernelGGL((batch_kernel<ST, 4>), grid, dim3(4 * kWave), 0, stream, items, segs, n_segs, counters, totals); break;
  }
  return hipGetLastError();
}

hipError_t launch_batch(int storage, bool staged, int stage_g, const MigArgs* items,
                        const BatchSeg* segs, int n_segs, unsigned* counters,
                        const unsigned* totals, int wpb, int blocks, hipStream_t stream) {
  switch (storage) {
    case PSM_STORE_F16: return launch_batch_st<PSM_STORE_F16>(staged, stage_g, items, segs, n_segs, counters, totals, wpb, blocks, stream);
    case PSM_STORE_BF16: return launch_batch_st<PSM_STORE_BF16>(staged, stage_g, items, segs, n_segs, counters, totals, wpb, blocks, stream);
    default: return launch_batch_st<PSM_STORE_F32>(staged, stage_g, items, segs, n_segs, counters, totals, wpb, blocks, stream);
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
