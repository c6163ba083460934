// velocity.cpp -- velocity providers turned into the Vrms cube on the data
// grid (VelocityProviderFactory / ConstantVelocityProvider /
// SegyVelocityProvider / TableVelocityProvider, velocity_reader.cpp:12-442,
// as prepareVelocitySliceWithPadding uses them, migration_kernel.cpp:470-511:
// one getVelocityIline per real inline).  The table provider's brute-force
// inverse-distance sum runs on the GPU (psm_table_velocity).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "host_internal.h"

namespace {

// VelocityProviderFactory::isNumeric (velocity_reader.cpp:373-397)
bool is_numeric(const std::string& s) {
  if (s.empty()) return false;
  size_t i = (s[0] == '-' || s[0] == '+') ? 1 : 0;
  bool dot = false;
  for (; i < s.size(); ++i) {
    if (s[i] == '.') {
      if (dot) return false;
      dot = true;
    } else if (!std::isdigit((unsigned char)s[i])) {
      return false;
    }
  }
  return true;
}

// VelocityProviderFactory::isSegyFile (:399-404)
bool is_segy(const std::string& p) {
  std::string l = p;
  std::transform(l.begin(), l.end(), l.begin(), [](unsigned char c) { return (char)std::tolower(c); });
  return l.find(".sgy") != std::string::npos || l.find(".segy") != std::string::npos;
}

// SegyVelocityProvider over the velocity file's own grid (:37-261).
struct SegyVel {
  psmh_segy* f = nullptr;
  // [n_il][n_xl][n_t] on the velocity grid (every element written by the
  // decode, so left uninitialised until then)
  std::unique_ptr<float[]> cube;
  size_t n_t = 0;
  double dt = 0.0;
  std::vector<double> times;
  ~SegyVel() { psmh_segy_close(f); }

  // readTraceByCoords: a missing trace reads as zeros (segy_utils.cpp:223-241)
  const float* trace(size_t a, size_t b) const { return cube.get() + (a * f->xl.size() + b) * n_t; }
  std::vector<float> zeros;

  // linear interpolation in time (:124-138, :183-196), double arithmetic
  float at_time(const float* tr, double time) const {
    if (time <= times[0]) return tr[0];
    if (time >= times[n_t - 1]) return tr[n_t - 1];
    const double q = time / dt;
    const size_t lo = (size_t)q;
    if (lo >= n_t - 1) return tr[n_t - 1];
    const double fr = q - (double)lo;
    return (float)(tr[lo] * (1.0 - fr) + tr[lo + 1] * fr);
  }

  // findNearestIndex (:70-91): ties go to the lower label
  static size_t nearest(const std::vector<int32_t>& v, int32_t x) {
    const auto it = std::lower_bound(v.begin(), v.end(), x);
    if (it == v.begin()) return 0;
    if (it == v.end()) return v.size() - 1;
    const size_t k = (size_t)(it - v.begin());
    return std::abs(v[k] - x) < std::abs(v[k - 1] - x) ? k : k - 1;
  }

  // interpolateVelocity (:93-163): the two neighbours (clamped) of the
  // nearest label on each axis, weights 1 / (1 + |dil| + |dxl|), float sums
  float interpolate(int32_t il, int32_t xl, double time) const {
    const std::vector<int32_t>& L = f->il;
    const std::vector<int32_t>& X = f->xl;
    const size_t a = nearest(L, il), b = nearest(X, xl);
    const size_t as[2] = {a > 0 ? a - 1 : 0, a < L.size() - 1 ? a + 1 : L.size() - 1};
    const size_t bs[2] = {b > 0 ? b - 1 : 0, b < X.size() - 1 ? b + 1 : X.size() - 1};
    float sv = 0.0f, sw = 0.0f;
    for (size_t ai : as)
      for (size_t bi : bs) {
        const float v = at_time(trace(ai, bi), time);
        const double d_il = std::abs((double)(L[ai] - il));
        const double d_xl = std::abs((double)(X[bi] - xl));
        const float w = 1.0f / (1.0f + (float)(d_il + d_xl));
        sv += v * w;
        sw += w;
      }
    return sw > 0 ? sv / sw : 0.0f;
  }

  // getVelocityIline (:203-261) for one data-grid inline
  void fill_inline(int32_t il, const int32_t* xls, size_t n_xl, size_t nt, double ddt, float* out) const {
    const auto ia = std::lower_bound(f->il.begin(), f->il.end(), il);
    const bool exact_il = ia != f->il.end() && *ia == il;
    for (size_t k = 0; k < n_xl; ++k) {
      float* o = out + k * nt;
      const int32_t xl = xls[k];
      const auto ib = std::lower_bound(f->xl.begin(), f->xl.end(), xl);
      const bool exact_xl = ib != f->xl.end() && *ib == xl;
      if (exact_il && exact_xl) {
        const float* tr = trace((size_t)(ia - f->il.begin()), (size_t)(ib - f->xl.begin()));
        if (std::abs(ddt - dt) < 1e-6 && nt == n_t) {
          std::copy(tr, tr + nt, o);
        } else {
          for (size_t t = 0; t < nt; ++t) o[t] = at_time(tr, (double)t * ddt);  // getVelocity, exact
        }
      } else {
        for (size_t t = 0; t < nt; ++t) o[t] = interpolate(il, xl, (double)t * ddt);
      }
    }
  }
};

void table_cube(psm_ctx* ctx, const std::string& path, const int32_t* il, size_t n_il,
                const int32_t* xl, size_t n_xl, size_t n_t, double dt, float* vel) {
  // TableVelocityProvider constructor (:264-318)
  std::ifstream f(path);
  if (!f.is_open()) throw std::runtime_error("Cannot open velocity table file: " + path);
  std::vector<psm_velocity_point> pts;
  std::string line;
  int n = 0;
  while (std::getline(f, line)) {
    ++n;
    if (line.empty() || line[0] == '#') continue;
    std::istringstream is(line);
    psm_velocity_point p;
    if (!(is >> p.inline_label >> p.crossline_label >> p.time >> p.velocity))
      throw std::runtime_error("Invalid format at line " + std::to_string(n) +
                               ": expected 4 columns (INLINE CROSSLINE TIME VEL)");
    pts.push_back(p);
  }
  if (pts.empty()) throw std::runtime_error("No valid data found in velocity table");
  if (!ctx) throw std::runtime_error("the table velocity provider needs a GPU context");
  const int rc = psm_table_velocity(ctx, pts.data(), pts.size(), il, n_il, xl, n_xl, n_t, dt, vel,
                                    PSM_HOST);
  if (rc != PSM_OK) throw std::runtime_error(psm_last_error(ctx));
}

void velocity_cube(psm_ctx* ctx, const char* param_c, const int32_t* il, size_t n_il,
                   const int32_t* xl, size_t n_xl, size_t n_t, double dt, float* vel) {
  if (!param_c || !il || !xl || !vel) throw std::runtime_error("null argument");
  const std::string param = param_c;
  const size_t n = n_il * n_xl * n_t;
  if (is_numeric(param)) {
    const float v = std::stof(param);
    if (v <= 0) throw std::runtime_error("Velocity must be positive, got: " + std::to_string(v));
    std::fill(vel, vel + n, v);
    return;
  }
  if (is_segy(param)) {
    SegyVel sv;
    int rc = psmh_segy_open(param.c_str(), &sv.f);
    if (rc != PSM_OK) throw std::runtime_error(psmh_last_error());
    sv.n_t = sv.f->n_t;
    sv.dt = sv.f->dt_us * 1e-6;
    sv.times.resize(sv.n_t);
    for (size_t i = 0; i < sv.n_t; ++i) sv.times[i] = (double)i * sv.dt;
    // the velocity file on the data's own grid (same labels, dt and sample
    // count): every trace is an exact copy (getVelocityIline's exact path,
    // a missing trace reading as zeros), so decode straight into vel
    if (sv.f->il.size() == n_il && sv.f->xl.size() == n_xl && sv.n_t == n_t &&
        std::abs(dt - sv.dt) < 1e-6 && std::equal(il, il + n_il, sv.f->il.begin()) &&
        std::equal(xl, xl + n_xl, sv.f->xl.begin())) {
      rc = psmh_segy_read_cube(sv.f, vel);
      if (rc != PSM_OK) throw std::runtime_error(psmh_last_error());
      return;
    }
    sv.cube.reset(new float[sv.f->il.size() * sv.f->xl.size() * sv.n_t]);
    rc = psmh_segy_read_cube(sv.f, sv.cube.get());
    if (rc != PSM_OK) throw std::runtime_error(psmh_last_error());
#pragma omp parallel for schedule(dynamic, 1)
    for (long long a = 0; a < (long long)n_il; ++a)
      sv.fill_inline(il[a], xl, n_xl, n_t, dt, vel + (size_t)a * n_xl * n_t);
    return;
  }
  table_cube(ctx, param, il, n_il, xl, n_xl, n_t, dt, vel);
}

}  // namespace

extern "C" int psmh_velocity_cube(psm_ctx* ctx, const char* param, const int32_t* il_labels,
                                  size_t n_il, const int32_t* xl_labels, size_t n_xl, size_t n_t,
                                  double dt, float* vel) {
  return psmh::guarded([&] { velocity_cube(ctx, param, il_labels, n_il, xl_labels, n_xl, n_t, dt, vel); });
}
