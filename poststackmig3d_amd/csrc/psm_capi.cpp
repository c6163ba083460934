// psm_capi.cpp -- the C ABI of libpsm_hip.so (declared in include/psm.h).
//
// Host side of the drop-in boundary: argument checks, the geometry of the
// reference driver (main.cpp:140-217, 258-288; migration_kernel.cpp:404-409),
// device buffer management and kernel dispatch.  No CPU fallback: every
// migration runs on the gfx950 kernels of psm_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "psm.h"
#include "psm_internal.h"

// Workspace of psm_migrate_volume_noalloc: the pair-layout cube, the per-trace
// scales, the output-inline tasks and the t0 table of one call, carved out of
// this array instead of allocated.  Each device holds its own copy for the
// life of the process.  Host code never uses the array's name as an address:
// its device address on each device comes from hipGetSymbolAddress
// (resolve_scratch) and is kept in g_scratch.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices];  // device address of `scratch` per device; nullptr: unresolved

// Looks up the device address of `scratch` on `device` (which must be the
// current device) the first time it is asked for that device.
hipError_t resolve_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e == hipSuccess) g_scratch[device] = p;
  return e;
}

void* scratch_address(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch[device];
}

// Grow-only device allocation owned by a context.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e == hipSuccess) cap = bytes;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

}  // namespace

struct psm_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  std::mutex mu;
  std::string err;
  DevBuf raw, pairs, scale, vel, out, t0, tasks, rowpos, counters, dot, dotx, doty, tvpts, tvlab;
  DevBuf adjrows, iorange, ioplane;  // staged adjoint: row-pair windows of a pass, their lists per data row
  int storage = PSM_STORE_F32;
  // kernel-level timing of the last volume call (psm_enable_timing)
  bool timing = false;
  bool timed = false;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  psm::Tuning tune;
  float alpha = 1.0f, alpha_mask = 1.0f;
};

// Device bytes the staged adjoint's row-pair buffer may take (passes of row
// offsets keep it within this; the 201 x 201 x 751 least-squares survey needs
// 4.9 GB for its 40 row offsets in one pass).
constexpr size_t kAdjRowsBytes = (size_t)8 << 30;

namespace {

int fail(psm_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define PSM_HIP(ctx, call)                                                           \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail((ctx), PSM_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define PSM_ALLOC(ctx, buf, bytes)                                                        \
  do {                                                                                    \
    hipError_t e_ = (buf).ensure(bytes);                                                  \
    if (e_ != hipSuccess)                                                                 \
      return fail((ctx), PSM_ERR_NOMEM, "device allocation of %zu bytes failed: %s",      \
                  (size_t)(bytes), hipGetErrorString(e_));                                \
  } while (0)

// ---------------------------------------------------------------- geometry
// main.cpp:144-172: labels extended by `pad` steps of the first-two-label step.
std::vector<int32_t> output_labels(const std::vector<int32_t>& labels, int pad) {
  if (pad > 0 && labels.size() > 1) {
    const int32_t step = labels[1] - labels[0];
    std::vector<int32_t> out;
    if (step <= 0) return out;
    const int32_t first = labels[0] - pad * step;
    const int32_t last = labels.back() + pad * step;
    for (int32_t l = first; l <= last; l += step) out.push_back(l);
    return out;
  }
  return labels;
}

// main.cpp:258-288: index of the real inline an output inline uses.
int32_t map_output_inline(const std::vector<int32_t>& labels, int32_t out_label) {
  for (size_t i = 0; i < labels.size(); ++i)
    if (labels[i] == out_label) return (int32_t)i;
  if (out_label < labels.front()) return 0;
  if (out_label > labels.back()) return (int32_t)labels.size() - 1;
  size_t best = 0;
  int32_t md = std::abs(out_label - labels[0]);
  for (size_t i = 1; i < labels.size(); ++i) {
    const int32_t d = std::abs(out_label - labels[i]);
    if (d < md) {
      md = d;
      best = i;
    }
  }
  return (int32_t)best;
}

std::vector<int32_t> labels_or_default(const int32_t* l, size_t n) {
  std::vector<int32_t> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = l ? l[i] : (int32_t)(i + 1);
  return v;
}

int check_survey(psm_ctx* c, const psm_survey* s) {
  if (!s) return fail(c, PSM_ERR_ARG, "survey is NULL");
  if (s->n_il == 0 || s->n_xl == 0 || s->n_t == 0)
    return fail(c, PSM_ERR_ARG, "empty survey (%zu x %zu x %zu)", s->n_il, s->n_xl, s->n_t);
  if (s->n_t > (1u << 20) || s->n_il > (1u << 20) || s->n_xl > (1u << 20))
    return fail(c, PSM_ERR_ARG, "survey dimension too large");
  if (!(s->dt > 0)) return fail(c, PSM_ERR_ARG, "dt must be positive, got %g", s->dt);
  if (!(s->inline_step > 0) || !(s->crossline_step > 0))
    return fail(c, PSM_ERR_ARG, "inline_step and crossline_step must be positive");
  if (s->inline_padding < 0 || s->crossline_padding < 0)
    return fail(c, PSM_ERR_ARG, "padding must be non-negative");
  if (!(s->angle_aperture > 0 && s->angle_aperture < 90))
    return fail(c, PSM_ERR_ARG, "angle_aperture must be in (0, 90), got %g", s->angle_aperture);
  for (int pass = 0; pass < 2; ++pass) {
    const int32_t* l = pass ? s->crossline_labels : s->inline_labels;
    const size_t n = pass ? s->n_xl : s->n_il;
    if (!l) continue;
    for (size_t i = 1; i < n; ++i)
      if (l[i] <= l[i - 1]) return fail(c, PSM_ERR_ARG, "labels must be strictly increasing");
  }
  return PSM_OK;
}

int plan_impl(psm_ctx* c, const psm_survey* s, const float* vel_probe, int where,
              psm_plan* plan, float alpha_mask = 1.0f) {
  int rc = check_survey(c, s);
  if (rc) return rc;
  if (!plan) return fail(c, PSM_ERR_ARG, "plan is NULL");
  std::memset(plan, 0, sizeof *plan);
  const std::vector<int32_t> ill = labels_or_default(s->inline_labels, s->n_il);
  const std::vector<int32_t> xll = labels_or_default(s->crossline_labels, s->n_xl);
  const std::vector<int32_t> oil = output_labels(ill, s->inline_padding);
  const std::vector<int32_t> oxl = output_labels(xll, s->crossline_padding);
  if (oil.empty() || oxl.empty()) return fail(c, PSM_ERR_ARG, "empty output geometry");

  // estimateMaxVelocity (main.cpp:22-64): 9 probes at the last sample
  float vmax = s->max_velocity;
  if (!(vmax > 0)) {
    if (!vel_probe) return fail(c, PSM_ERR_ARG, "max_velocity <= 0 and no velocity cube to probe");
    std::vector<size_t> ils{0}, xls{0};
    if (s->n_il > 1) { ils.push_back(s->n_il / 2); ils.push_back(s->n_il - 1); }
    if (s->n_xl > 1) { xls.push_back(s->n_xl / 2); xls.push_back(s->n_xl - 1); }
    vmax = 0.0f;
    for (size_t a : ils)
      for (size_t b : xls) {
        const size_t off = (a * s->n_xl + b) * s->n_t + (s->n_t - 1);
        float v = 0.0f;
        if (where == PSM_DEVICE) {
          // on the context's stream, so the probe sees what earlier work on
          // that stream (or work it was made to wait for) wrote into vel
          hipError_t e = c ? hipMemcpyAsync(&v, vel_probe + off, sizeof(float),
                                            hipMemcpyDeviceToHost, c->stream)
                           : hipMemcpy(&v, vel_probe + off, sizeof(float), hipMemcpyDeviceToHost);
          if (e == hipSuccess && c) e = hipStreamSynchronize(c->stream);
          if (e != hipSuccess)
            return fail(c, PSM_ERR_HIP, "velocity probe copy failed: %s", hipGetErrorString(e));
        } else {
          v = vel_probe[off];
        }
        if (v > vmax) vmax = v;
      }
    if (vmax < 100.0f) vmax = 3000.0f;
  }
  // a scaled velocity model scales the probed maximum (exact for the
  // reference operator, alpha_mask = 1)
  vmax *= alpha_mask;
  // main.cpp:188-202
  const float tan_theta = (float)std::tan(s->angle_aperture * 3.14159265358979323846 / 180.0);
  const float tan2 = tan_theta * tan_theta;
  const double max_t = (double)(s->n_t - 1) * s->dt;
  const float radius = (vmax * (float)max_t / 2.0f) * std::sqrt(tan2);
  const int ap = (int)((double)radius / s->inline_step) + 1;

  plan->n_out_il = oil.size();
  plan->n_out_xl = oxl.size();
  plan->n_t = s->n_t;
  plan->max_aperture_il = ap;
  plan->tan_theta_sq = tan2;
  plan->max_velocity = vmax;
  plan->out_inline_labels = (int32_t*)std::malloc(oil.size() * sizeof(int32_t));
  plan->out_crossline_labels = (int32_t*)std::malloc(oxl.size() * sizeof(int32_t));
  plan->out_to_in_il = (int32_t*)std::malloc(oil.size() * sizeof(int32_t));
  plan->window = (int32_t*)std::malloc(2 * oil.size() * sizeof(int32_t));
  plan->t0_times = (float*)std::malloc(s->n_t * sizeof(float));
  if (!plan->out_inline_labels || !plan->out_crossline_labels || !plan->out_to_in_il ||
      !plan->window || !plan->t0_times) {
    psm_plan_free(plan);
    return fail(c, PSM_ERR_NOMEM, "host allocation failed");
  }
  std::memcpy(plan->out_inline_labels, oil.data(), oil.size() * sizeof(int32_t));
  std::memcpy(plan->out_crossline_labels, oxl.data(), oxl.size() * sizeof(int32_t));
  const int n_il = (int)s->n_il;
  for (size_t io = 0; io < oil.size(); ++io) {
    const int cidx = map_output_inline(ill, oil[io]);
    plan->out_to_in_il[io] = cidx;
    // migration_kernel.cpp:404-409
    const int s_req = std::max(0, cidx - ap);
    const int e_req = std::min(n_il, cidx + ap + 1);
    plan->window[2 * io] = std::max(0, s_req - s->inline_padding);
    plan->window[2 * io + 1] = std::min(n_il, e_req + s->inline_padding);
  }
  for (size_t i = 0; i < s->n_t; ++i) plan->t0_times[i] = (float)((double)i * s->dt);  // main.cpp:216
  return PSM_OK;
}

hipError_t upload(DevBuf& b, const void* src, size_t bytes, hipStream_t st) {
  hipError_t e = b.ensure(bytes);
  if (e != hipSuccess) return e;
  return hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, st);
}

}  // namespace

extern "C" {

const char* psm_version(void) { return "poststackmig3d_amd 0.1 (gfx950)"; }

int psm_create(psm_ctx** out, int device) {
  if (!out) return PSM_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return PSM_ERR_NODEV;
  if (device < 0 || device >= n) return PSM_ERR_NODEV;
  psm_ctx* c = new (std::nothrow) psm_ctx();
  if (!c) return PSM_ERR_NOMEM;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return PSM_ERR_HIP;
  }
  c->stream = c->own_stream;
  // once per device: the device address of the no-allocation workspace (a
  // failure only disables psm_migrate_volume_noalloc, which reports it)
  (void)resolve_scratch(device);
  *out = c;
  return PSM_OK;
}

void psm_destroy(psm_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (DevBuf* b : {&c->raw, &c->pairs, &c->scale, &c->vel, &c->out, &c->t0, &c->tasks,
                    &c->rowpos, &c->counters, &c->dot, &c->dotx, &c->doty, &c->tvpts, &c->tvlab,
                    &c->adjrows, &c->iorange, &c->ioplane})
    b->release();
  for (hipEvent_t& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
}

const char* psm_last_error(const psm_ctx* c) { return c ? c->err.c_str() : "no context"; }

int psm_set_stream(psm_ctx* c, void* stream) {
  if (!c) return PSM_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  c->stream = static_cast<hipStream_t>(stream);
  return PSM_OK;
}

int psm_get_stream(psm_ctx* c, void** stream) {
  if (!c || !stream) return PSM_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  *stream = static_cast<void*>(c->stream);
  return PSM_OK;
}

int psm_synchronize(psm_ctx* c) {
  if (!c) return PSM_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  PSM_HIP(c, hipSetDevice(c->device));
  PSM_HIP(c, hipStreamSynchronize(c->stream));
  return PSM_OK;
}

int psm_kirchhoff_inline(psm_ctx* c, const float* input_buffer, size_t n_buf_il,
                         size_t n_buf_xl, size_t n_t, const float* velocity_slice,
                         float* output_slice, size_t n_xl, const int32_t* il_pos,
                         size_t n_il_pos, const int32_t* xl_pos, size_t n_xl_pos,
                         int32_t current_il_pos, int32_t current_xl_start, float dt,
                         const float* t0_times, float tan_theta_sq, float inline_bin_size,
                         float crossline_bin_size, int amp_correction) {
  (void)xl_pos;
  (void)n_xl_pos;
  if (!c) return PSM_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  if (n_xl == 0 || n_t == 0) return PSM_OK;  // migration_kernel.cpp:35-37
  if (!output_slice || !t0_times || !velocity_slice)
    return fail(c, PSM_ERR_ARG, "null output, velocity or t0 pointer");
  if (n_buf_il > 0 && (!input_buffer || !il_pos))
    return fail(c, PSM_ERR_ARG, "null input buffer or il_pos_indices");
  if (n_il_pos < n_buf_il)
    return fail(c, PSM_ERR_ARG, "il_pos_indices has %zu entries, need %zu", n_il_pos, n_buf_il);
  if (n_t > (1u << 24) || n_buf_xl > (1u << 24) || n_xl > (1u << 24) || n_buf_il > (1u << 24))
    return fail(c, PSM_ERR_ARG, "dimension too large");
  if (!(dt > 0) || !(crossline_bin_size > 0))
    return fail(c, PSM_ERR_ARG, "dt and crossline_bin_size must be positive");
  PSM_HIP(c, hipSetDevice(c->device));
  hipStream_t st = c->stream;

  const int32_t cpos = (current_il_pos >= 0 && (size_t)current_il_pos < n_il_pos)
                           ? il_pos[current_il_pos] : 0;  // migration_kernel.cpp:48-49
  bool contiguous = n_buf_il > 0;
  for (size_t b = 1; b < n_buf_il && contiguous; ++b)
    contiguous = il_pos[b] == il_pos[0] + (int32_t)b;
  long long crow = n_buf_il ? (long long)cpos - il_pos[0] : 0;
  if (std::llabs(crow) > (1LL << 28)) contiguous = false;

  const size_t n_traces = n_buf_il * n_buf_xl;
  const size_t in_bytes = n_traces * n_t * sizeof(float);
  if ((n_buf_xl + 1) * n_t * sizeof(float2) >= (size_t)INT32_MAX)
    return fail(c, PSM_ERR_ARG, "one inline of the buffer exceeds 2 GiB");
  PSM_ALLOC(c, c->pairs, psm::pairs_bytes(c->storage, (int)n_buf_il, (int)n_buf_xl, (int)n_t));
  PSM_ALLOC(c, c->scale, (n_buf_il + 1) * (n_buf_xl + 1) * sizeof(float));
  if (n_traces) PSM_HIP(c, upload(c->raw, input_buffer, in_bytes, st));
  PSM_HIP(c, psm::launch_pack(c->storage, c->raw.as<float>(), c->pairs.p, c->scale.as<float>(),
                              (int)n_buf_il, (int)n_buf_xl, (int)n_t, st));
  PSM_HIP(c, upload(c->vel, velocity_slice, n_buf_xl * n_t * sizeof(float), st));
  PSM_HIP(c, upload(c->t0, t0_times, n_t * sizeof(float), st));
  psm::Task tk{contiguous ? (int)crow : 0, 0, (int)n_buf_il, 0};
  PSM_HIP(c, upload(c->tasks, &tk, sizeof tk, st));
  if (!contiguous && n_buf_il) PSM_HIP(c, upload(c->rowpos, il_pos, n_buf_il * sizeof(int32_t), st));
  PSM_ALLOC(c, c->out, n_xl * n_t * sizeof(float));

  psm::MigArgs a{};
  a.xp = c->pairs.p;
  a.scale = c->scale.as<float>();
  a.n_rows = (int)n_buf_il;
  a.storage = c->storage;
  a.vel = c->vel.as<float>();
  a.out = c->out.as<float>();
  a.t0 = c->t0.as<float>();
  a.tasks = c->tasks.as<psm::Task>();
  a.row_pos = (!contiguous && n_buf_il) ? c->rowpos.as<int>() : nullptr;
  a.counters = nullptr;
  a.n_t = (int)n_t;
  a.n_xd = (int)n_buf_xl;
  a.xd_off = 0;
  a.n_xv = (int)n_buf_xl;
  a.xv_off = 0;
  a.K = (int)n_buf_xl;
  a.n_out_xl = (int)n_xl;
  a.cur_xl_start = current_xl_start;
  a.n_out_il = 1;
  a.cpos = cpos;
  a.inv_dt = 1.0f / dt;                      // migration_kernel.cpp:44
  a.dx = inline_bin_size;
  a.dy = crossline_bin_size;
  a.inv_dy = 1.0f / crossline_bin_size;      // :45
  a.tan2 = tan_theta_sq;
  a.amp = amp_correction ? 1 : 0;
  a.alpha = c->alpha;
  a.alpha_mask = c->alpha_mask;
  if (n_buf_il == 0) {
    // no aperture traces: every valid sample sums to zero
    a.tasks = c->tasks.as<psm::Task>();
  }
  int extent = 0;
  if (a.row_pos)
    extent = (int)n_buf_il - 1;
  else if (n_buf_il)
    extent = (int)std::max<long long>(0, std::max<long long>((long long)n_buf_il - 1 - crow, crow));
  PSM_HIP(c, psm::launch_migrate(a, psm::kForward, st, c->tune, extent));
  PSM_HIP(c, hipMemcpyAsync(output_slice, c->out.p, n_xl * n_t * sizeof(float), hipMemcpyDeviceToHost, st));
  PSM_HIP(c, hipStreamSynchronize(st));
  return PSM_OK;
}

int psm_survey_plan(const psm_survey* s, const float* vel_probe, int where, psm_plan* plan) {
  return plan_impl(nullptr, s, vel_probe, where, plan);
}

void psm_plan_free(psm_plan* p) {
  if (!p) return;
  std::free(p->out_inline_labels);
  std::free(p->out_crossline_labels);
  std::free(p->out_to_in_il);
  std::free(p->window);
  std::free(p->t0_times);
  std::memset(p, 0, sizeof *p);
}

}  // extern "C"

namespace {

// Shared body of the volume entry points (lock held).  mode:
//   kForward  in = data cube,  out = image
//   kTangent  in = data cube,  out = d image / d alpha
//   kAdjoint  in = image,      out = data cube
//   kCount    in unused,       out unused; counts (and count images)
int volume_impl(psm_ctx* c, const psm_survey* s, const float* in, const float* vel, float* out,
                int where, int mode, double* p_ap = nullptr, double* p_g = nullptr,
                float* ap_image = nullptr, float* g_image = nullptr, bool noalloc = false) {
  const bool count = mode == psm::kCount;
  const bool adjoint = mode == psm::kAdjoint;
  const bool gathers = mode == psm::kForward || mode == psm::kTangent;
  if (where != PSM_HOST && where != PSM_DEVICE) return fail(c, PSM_ERR_ARG, "bad `where`");
  if (!vel || (!count && (!in || !out))) return fail(c, PSM_ERR_ARG, "null input/vel/out pointer");
  PSM_HIP(c, hipSetDevice(c->device));
  psm_plan plan;
  int rc = plan_impl(c, s, vel, where, &plan, c->alpha_mask);
  if (rc) return rc;
  struct PlanGuard {
    psm_plan* p;
    ~PlanGuard() { psm_plan_free(p); }
  } guard{&plan};
  hipStream_t st = c->stream;
  const size_t n_tr = s->n_il * s->n_xl, nt = s->n_t;
  const size_t cube = n_tr * nt;
  const size_t out_elems = plan.n_out_il * plan.n_out_xl * nt;
  const int padx = s->crossline_padding;
  const size_t K = s->n_xl + 2 * (size_t)padx;
  const int storage = mode == psm::kForward ? c->storage : PSM_STORE_F32;

  std::vector<psm::Task> tasks(plan.n_out_il);
  int extent = 0;
  for (size_t io = 0; io < plan.n_out_il; ++io) {
    tasks[io] = psm::Task{plan.out_to_in_il[io], plan.window[2 * io], plan.window[2 * io + 1],
                          plan.out_to_in_il[io]};
    extent = std::max(extent, std::max(tasks[io].b1 - 1 - tasks[io].c, tasks[io].c - tasks[io].b0));
  }
  if (((size_t)s->n_xl + 1) * nt * sizeof(float2) >= (size_t)INT32_MAX)
    return fail(c, PSM_ERR_ARG, "one inline of the survey exceeds 2 GiB");
  const size_t pairs_bytes = psm::pairs_bytes(storage, (int)s->n_il, (int)s->n_xl, (int)nt);
  const size_t scale_bytes = (s->n_il + 1) * (s->n_xl + 1) * sizeof(float);
  const size_t tasks_bytes = tasks.size() * sizeof(psm::Task);

  // Workspace: the context's grow-only buffers, or (noalloc) consecutive
  // 256-byte aligned pieces of the static `scratch` array at this device's
  // address: tasks, t0 table, scales, pairs.
  psm::Task* w_tasks;
  float* w_t0;
  float* w_scale = nullptr;
  void* w_pairs = nullptr;
  if (noalloc) {
    char* const base = static_cast<char*>(scratch_address(c->device));
    if (!base)
      return fail(c, PSM_ERR_HIP, "the no-allocation workspace has no device address on device %d",
                  c->device);
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t need = up(tasks_bytes) + up(nt * sizeof(float)) + up(scale_bytes) + up(pairs_bytes);
    if (need > kScratchBytes)
      return fail(c, PSM_ERR_ARG,
                  "survey %zu x %zu x %zu needs %zu bytes of workspace; "
                  "psm_migrate_volume_noalloc holds at most %zu",
                  s->n_il, s->n_xl, nt, need, kScratchBytes);
    w_tasks = reinterpret_cast<psm::Task*>(base);
    w_t0 = reinterpret_cast<float*>(base + up(tasks_bytes));
    w_scale = reinterpret_cast<float*>(base + up(tasks_bytes) + up(nt * sizeof(float)));
    w_pairs = base + up(tasks_bytes) + up(nt * sizeof(float)) + up(scale_bytes);
    PSM_HIP(c, hipMemcpyAsync(w_tasks, tasks.data(), tasks_bytes, hipMemcpyHostToDevice, st));
    PSM_HIP(c, hipMemcpyAsync(w_t0, plan.t0_times, nt * sizeof(float), hipMemcpyHostToDevice, st));
  } else {
    PSM_HIP(c, upload(c->tasks, tasks.data(), tasks_bytes, st));
    PSM_HIP(c, upload(c->t0, plan.t0_times, nt * sizeof(float), st));
    w_tasks = c->tasks.as<psm::Task>();
    w_t0 = c->t0.as<float>();
  }

  // device views of the inputs and outputs
  const float* d_vel = vel;
  const float* d_in = in;
  float* d_out = out;
  const size_t in_elems = adjoint ? out_elems : cube;
  const size_t res_elems = adjoint ? cube : out_elems;
  if (where == PSM_HOST) {
    PSM_HIP(c, upload(c->vel, vel, cube * sizeof(float), st));
    d_vel = c->vel.as<float>();
    if (!count) {
      PSM_HIP(c, upload(c->raw, in, in_elems * sizeof(float), st));
      d_in = c->raw.as<float>();
      PSM_ALLOC(c, c->out, res_elems * sizeof(float));
      d_out = c->out.as<float>();
    }
  }
  if (gathers && !noalloc) {
    PSM_ALLOC(c, c->pairs, pairs_bytes);
    PSM_ALLOC(c, c->scale, scale_bytes);
    w_pairs = c->pairs.p;
    w_scale = c->scale.as<float>();
  }
  const bool timed = c->timing && mode == psm::kForward && !noalloc;
  c->timed = false;
  if (timed) {
    for (hipEvent_t& e : c->ev)
      if (!e) PSM_HIP(c, hipEventCreate(&e));
    PSM_HIP(c, hipEventRecord(c->ev[0], st));
  }
  if (gathers)
    PSM_HIP(c, psm::launch_pack(storage, d_in, w_pairs, w_scale, (int)s->n_il, (int)s->n_xl,
                                (int)nt, st));
  if (timed) PSM_HIP(c, hipEventRecord(c->ev[1], st));

  psm::MigArgs a{};
  a.xp = w_pairs;
  a.scale = w_scale;
  a.n_rows = (int)s->n_il;
  a.storage = storage;
  a.vel = d_vel;
  a.out = gathers ? d_out : nullptr;
  a.t0 = w_t0;
  a.tasks = w_tasks;
  a.row_pos = nullptr;
  a.n_t = (int)nt;
  a.n_xd = (int)s->n_xl;
  a.xd_off = padx;
  a.n_xv = (int)s->n_xl;
  a.xv_off = padx;
  a.K = (int)K;
  a.n_out_xl = (int)plan.n_out_xl;
  a.cur_xl_start = 0;  // main.cpp:362
  a.n_out_il = (int)plan.n_out_il;
  a.cpos = 0;
  const float dtf = (float)s->dt;  // main.cpp:363
  a.inv_dt = 1.0f / dtf;
  a.dx = (float)s->inline_step;
  a.dy = (float)s->crossline_step;
  a.inv_dy = 1.0f / a.dy;
  a.tan2 = plan.tan_theta_sq;
  a.amp = s->amp_correction ? 1 : 0;
  a.alpha = c->alpha;
  a.alpha_mask = c->alpha_mask;
  if (adjoint) {
    a.image = d_in;
    a.adj_out = d_out;
    PSM_HIP(c, psm::launch_fill_zero(d_out, cube, st));
  }
  if (count) {
    PSM_ALLOC(c, c->counters, 2 * sizeof(unsigned long long));
    PSM_HIP(c, hipMemsetAsync(c->counters.p, 0, 2 * sizeof(unsigned long long), st));
    a.counters = c->counters.as<unsigned long long>();
    a.out = nullptr;
    a.out_g = nullptr;
    if (ap_image || g_image) {
      if (where == PSM_DEVICE) {
        a.out = ap_image;
        a.out_g = g_image;
      } else {
        PSM_ALLOC(c, c->out, 2 * out_elems * sizeof(float));
        a.out = ap_image ? c->out.as<float>() : nullptr;
        a.out_g = g_image ? c->out.as<float>() + out_elems : nullptr;
      }
    }
    PSM_HIP(c, psm::launch_migrate(a, psm::kCount, st, c->tune, extent));
    unsigned long long h[2] = {0, 0};
    PSM_HIP(c, hipMemcpyAsync(h, c->counters.p, sizeof h, hipMemcpyDeviceToHost, st));
    if (where == PSM_HOST) {
      if (ap_image)
        PSM_HIP(c, hipMemcpyAsync(ap_image, a.out, out_elems * sizeof(float), hipMemcpyDeviceToHost, st));
      if (g_image)
        PSM_HIP(c, hipMemcpyAsync(g_image, a.out_g, out_elems * sizeof(float), hipMemcpyDeviceToHost, st));
    }
    PSM_HIP(c, hipStreamSynchronize(st));
    if (p_ap) *p_ap = (double)h[0];
    if (p_g) *p_g = (double)h[1];
    return PSM_OK;
  }
  // Staged adjoint: each column window is added once into a row-pair buffer
  // [n_out_il][nd][n_xl][n_t] (row offsets d - d_lo of one pass) instead of
  // into both data rows c + d and c - d, and a second kernel adds the buffer
  // into the data rows (half the global atomics); passes of nd row offsets
  // keep the buffer within kAdjRowsBytes.  rows_per_pass > 0 sets nd.
  const size_t per_d = plan.n_out_il * (size_t)s->n_xl * nt;  // floats per row offset
  int nd = 0;
  if (adjoint && psm::adjoint_staged(a, c->tune) && per_d > 0) {
    // at most kAdjRowsBytes and half of the free device memory (beyond what
    // the context holds already)
    size_t free_b = 0, total_b = 0;
    size_t budget = kAdjRowsBytes;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
      budget = std::min(budget, (free_b + c->adjrows.cap) / 2);
    const size_t fit = budget / (per_d * sizeof(float));
    nd = (int)std::min<size_t>(fit, (size_t)extent + 1);
    if (c->tune.rows_per_pass > 0) nd = std::min(nd, c->tune.rows_per_pass);
    // no buffer: the windows are added into both data rows directly
    if (nd > 0 && c->adjrows.ensure((size_t)nd * per_d * sizeof(float)) != hipSuccess) {
      (void)hipGetLastError();
      nd = 0;
    }
  }
  if (nd > 0) {
    a.adj_rows = c->adjrows.as<float>();
    a.adj_nd = nd;
    // per pass and data row r, the planes io * nd + (d - d0) of the
    // row-pair buffer whose pair holds r: output inlines io with
    // b0 <= r < b1 and d = |r - c(io)| in the pass, in output-inline order;
    // all passes' lists are uploaded once, before the first launch
    const size_t n_rows = s->n_il;
    const int n_pass = extent / nd + 1;
    std::vector<int> start((size_t)n_pass * (n_rows + 1)), plane;
    for (int p = 0; p < n_pass; ++p) {
      const int d0 = p * nd, d1 = std::min(d0 + nd, extent + 1);
      int* const st_p = start.data() + (size_t)p * (n_rows + 1);
      for (size_t r = 0; r < n_rows; ++r) {
        st_p[r] = (int)plane.size();
        for (size_t io = 0; io < tasks.size(); ++io) {
          const psm::Task& tk = tasks[io];
          const int d = std::abs((int)r - tk.c);
          if ((int)r >= tk.b0 && (int)r < tk.b1 && d >= d0 && d < d1)
            plane.push_back((int)io * nd + (d - d0));
        }
      }
      st_p[n_rows] = (int)plane.size();
    }
    PSM_HIP(c, upload(c->iorange, start.data(), start.size() * sizeof(int), st));
    PSM_HIP(c, upload(c->ioplane, plane.data(), std::max<size_t>(plane.size(), 1) * sizeof(int), st));
    for (int p = 0; p < n_pass; ++p) {
      const int d0 = p * nd;
      PSM_HIP(c, psm::launch_fill_zero(a.adj_rows, (size_t)nd * per_d, st));
      a.d_lo = d0;
      a.d_hi = std::min(d0 + nd, extent + 1);
      a.accumulate = d0 > 0 ? 1 : 0;
      PSM_HIP(c, psm::launch_adjoint_staged(a, st));
      PSM_HIP(c, psm::launch_adjoint_rows(a.adj_rows, d_out,
                                          c->iorange.as<int>() + (size_t)p * (n_rows + 1),
                                          c->ioplane.as<int>(), (int)n_rows, (int)s->n_xl,
                                          (int)nt, st));
    }
  } else {
    PSM_HIP(c, psm::launch_migrate(a, mode, st, c->tune, extent));
  }
  if (timed) {
    PSM_HIP(c, hipEventRecord(c->ev[2], st));
    c->timed = true;
  }
  if (where == PSM_HOST) {
    PSM_HIP(c, hipMemcpyAsync(out, d_out, res_elems * sizeof(float), hipMemcpyDeviceToHost, st));
    PSM_HIP(c, hipStreamSynchronize(st));
  }
  return PSM_OK;
}

}  // namespace

extern "C" {

int psm_migrate_volume(psm_ctx* c, const psm_survey* s, const float* data, const float* vel,
                       float* out, int where) {
  if (!c) return PSM_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  return volume_impl(c, s, data, vel, out, where, psm::kForward);
}

int psm_migrate_volume_noalloc(psm_ctx* c, const psm_survey* s, const float* data,
                               const float* vel, float* out) {
  if (!c) return PSM_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  return volume_impl(c, s, data, vel, out, PSM_DEVICE, psm::kForward, nullptr, nullptr, nullptr,
                     nullptr, true);
}

int psm_demigrate_volume(psm_ctx* c, const psm_survey* s, const float* image, const float* vel,
                         float* data, int where) {
  if (!c) return PSM_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  return volume_impl(c, s, image, vel, data, where, psm::kAdjoint);
}

int psm_migrate_volume_dalpha(psm_ctx* c, const psm_survey* s, const float* data,
                              const float* vel, float* out, int where) {
  if (!c) return PSM_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  return volume_impl(c, s, data, vel, out, where, psm::kTangent);
}

int psm_dot_f64(psm_ctx* c, const float* x, const float* y, size_t n, int where, double* result) {
  if (!c || !result) return PSM_ERR_ARG;
  if (n && (!x || !y)) return fail(c, PSM_ERR_ARG, "null vector");
  if (where != PSM_HOST && where != PSM_DEVICE) return fail(c, PSM_ERR_ARG, "bad `where`");
  std::lock_guard<std::mutex> lk(c->mu);
  PSM_HIP(c, hipSetDevice(c->device));
  hipStream_t st = c->stream;
  const float* dx = x;
  const float* dy = y;
  if (where == PSM_HOST && n) {
    PSM_HIP(c, upload(c->dotx, x, n * sizeof(float), st));
    PSM_HIP(c, upload(c->doty, y, n * sizeof(float), st));
    dx = c->dotx.as<float>();
    dy = c->doty.as<float>();
  }
  PSM_ALLOC(c, c->dot, psm::dot_scratch_doubles() * sizeof(double));
  double* part = c->dot.as<double>();
  PSM_HIP(c, psm::launch_dot_f64(dx, dy, n, part, part + psm::dot_scratch_doubles() - 1, st));
  PSM_HIP(c, hipMemcpyAsync(result, part + psm::dot_scratch_doubles() - 1, sizeof(double),
                            hipMemcpyDeviceToHost, st));
  PSM_HIP(c, hipStreamSynchronize(st));
  return PSM_OK;
}

int psm_table_velocity(psm_ctx* c, const psm_velocity_point* points, size_t n_points,
                       const int32_t* il, size_t n_il, const int32_t* xl, size_t n_xl, size_t n_t,
                       double dt, float* vel, int where) {
  if (!c) return PSM_ERR_ARG;
  if (!points || n_points == 0) return fail(c, PSM_ERR_ARG, "No valid data found in velocity table");
  if (!il || !xl || !vel) return fail(c, PSM_ERR_ARG, "null labels or output");
  if (where != PSM_HOST && where != PSM_DEVICE) return fail(c, PSM_ERR_ARG, "bad `where`");
  if (n_points > (size_t)INT32_MAX || n_il > (1u << 20) || n_xl > (1u << 20) || n_t > (1u << 20))
    return fail(c, PSM_ERR_ARG, "table velocity: size out of range");
  std::lock_guard<std::mutex> lk(c->mu);
  PSM_HIP(c, hipSetDevice(c->device));
  hipStream_t st = c->stream;
  float vmin = points[0].velocity;
  for (size_t k = 1; k < n_points; ++k) vmin = std::min(vmin, points[k].velocity);
  PSM_HIP(c, upload(c->tvpts, points, n_points * sizeof(psm_velocity_point), st));
  std::vector<int32_t> lab(il, il + n_il);
  lab.insert(lab.end(), xl, xl + n_xl);
  PSM_HIP(c, upload(c->tvlab, lab.data(), lab.size() * sizeof(int32_t), st));
  const size_t n = n_il * n_xl * n_t;
  float* dvel = vel;
  if (where == PSM_HOST) {
    PSM_ALLOC(c, c->out, n * sizeof(float));
    dvel = c->out.as<float>();
  }
  const int32_t* dl = c->tvlab.as<int32_t>();
  PSM_HIP(c, psm::launch_table_velocity(c->tvpts.as<psm_velocity_point>(), (int)n_points, vmin, dl,
                                        dl + n_il, (int)n_il, (int)n_xl, (int)n_t, dt, dvel, st));
  if (where == PSM_HOST)
    PSM_HIP(c, hipMemcpyAsync(vel, dvel, n * sizeof(float), hipMemcpyDeviceToHost, st));
  PSM_HIP(c, hipStreamSynchronize(st));
  return PSM_OK;
}

int psm_alloc_host(size_t bytes, void** ptr) {
  if (!ptr) return PSM_ERR_ARG;
  *ptr = nullptr;
  if (hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
    *ptr = nullptr;
    return PSM_ERR_NOMEM;
  }
  return PSM_OK;
}

void psm_free_host(void* ptr) {
  if (ptr) (void)hipHostFree(ptr);
}

int psm_set_velocity_scale(psm_ctx* c, float alpha, float alpha_mask) {
  if (!c) return PSM_ERR_ARG;
  if (!(alpha > 0) || !(alpha_mask > 0)) return fail(c, PSM_ERR_ARG, "velocity scales must be > 0");
  std::lock_guard<std::mutex> lk(c->mu);
  c->alpha = alpha;
  c->alpha_mask = alpha_mask;
  return PSM_OK;
}

int psm_count_pairs(psm_ctx* c, const psm_survey* s, const float* vel, int where, double* p_ap,
                    double* p_g, float* ap_image, float* g_image) {
  if (!c) return PSM_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  return volume_impl(c, s, nullptr, vel, nullptr, where, psm::kCount, p_ap, p_g, ap_image, g_image);
}

int psm_set_storage(psm_ctx* c, int storage) {
  if (!c) return PSM_ERR_ARG;
  if (storage != PSM_STORE_F32 && storage != PSM_STORE_F16 && storage != PSM_STORE_BF16)
    return fail(c, PSM_ERR_ARG, "unknown storage %d", storage);
  std::lock_guard<std::mutex> lk(c->mu);
  c->storage = storage;
  return PSM_OK;
}

int psm_set_tuning(psm_ctx* c, int waves_per_block, int rows_per_pass) {
  if (!c) return PSM_ERR_ARG;
  if (waves_per_block != 1 && waves_per_block != 2 && waves_per_block != 4 &&
      waves_per_block != 8 && waves_per_block != 16)
    return fail(c, PSM_ERR_ARG, "waves_per_block must be 1, 2, 4, 8 or 16");
  if (rows_per_pass < -1) return fail(c, PSM_ERR_ARG, "rows_per_pass must be >= -1");
  std::lock_guard<std::mutex> lk(c->mu);
  c->tune.waves_per_block = waves_per_block;
  c->tune.rows_per_pass = rows_per_pass;
  return PSM_OK;
}

int psm_get_tuning(psm_ctx* c, int* waves_per_block, int* rows_per_pass) {
  if (!c) return PSM_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  if (waves_per_block) *waves_per_block = c->tune.waves_per_block;
  if (rows_per_pass) *rows_per_pass = c->tune.rows_per_pass;
  return PSM_OK;
}

int psm_set_staging(psm_ctx* c, int on, int steps) {
  if (!c) return PSM_ERR_ARG;
  if (on < -1 || on > 1) return fail(c, PSM_ERR_ARG, "staging must be -1, 0 or 1");
  if (steps != 2 && steps != 4) return fail(c, PSM_ERR_ARG, "steps must be 2 or 4");
  std::lock_guard<std::mutex> lk(c->mu);
  c->tune.staging = on;
  c->tune.stage_steps = steps;
  return PSM_OK;
}

int psm_get_staging(psm_ctx* c, int* on, int* steps) {
  if (!c) return PSM_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  if (on) *on = c->tune.staging;
  if (steps) *steps = c->tune.stage_steps;
  return PSM_OK;
}

int psm_enable_timing(psm_ctx* c, int on) {
  if (!c) return PSM_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  c->timing = on != 0;
  return PSM_OK;
}

int psm_last_timing(psm_ctx* c, psm_timing* t) {
  if (!c || !t) return PSM_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  if (!c->timed) return fail(c, PSM_ERR_ARG, "no timed volume call on this context");
  PSM_HIP(c, hipSetDevice(c->device));
  PSM_HIP(c, hipEventSynchronize(c->ev[2]));
  PSM_HIP(c, hipEventElapsedTime(&t->pack_ms, c->ev[0], c->ev[1]));
  PSM_HIP(c, hipEventElapsedTime(&t->migrate_ms, c->ev[1], c->ev[2]));
  return PSM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- batches
// Resident workgroups per CU of the staged batch form by default: 8-wave
// workgroups are held to three per CU by their registers (80 per lane at six
// waves per SIMD), 4-wave workgroups to five by their LDS (about 30 KB each
// with 4 offsets per group).
constexpr int kBatchStagedPerCu = 3;
constexpr int kBatchStaged4PerCu = 5;
// Several volumes (a velocity scan over scaled Vrms models, sub-volumes of
// any size) migrated by one persistent launch, captured once into a HIP graph
// together with the trace packing and the queue reset, then replayed.

struct psm_batch {
  psm_ctx* ctx = nullptr;
  int storage = PSM_STORE_F32;
  int n_items = 0;
  int wpb = 4, blocks_per_cu = 0;
  bool staged = false;  // kernel form of the captured launch
  int stage_g = psm::kAutoStageSteps;  // staged form: crossline offsets per staged group
  int n_cu = 0;
  std::vector<psm::MigArgs> args;
  std::vector<psm::BatchSeg> segs;
  unsigned totals[psm::kBatchXcds] = {};  // items per XCD queue
  // per distinct data cube: raw pointer, shape, packed buffer, scales
  struct Group {
    const float* data;
    int n_il, n_xl, n_t;
    DevBuf pairs, scale;
  };
  std::vector<std::unique_ptr<Group>> groups;
  std::vector<std::unique_ptr<DevBuf>> item_bufs;  // tasks and t0 per item
  DevBuf d_args, d_segs, d_totals, counter;
  std::vector<int> n_tiles;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

void batch_release_graph(psm_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Captures [queue reset, packing of every data cube, persistent kernel].
int batch_capture(psm_batch* b) {
  psm_ctx* c = b->ctx;
  batch_release_graph(b);
  hipStream_t st = c->own_stream;  // capture needs a stream of our own
  const int blocks = b->n_cu * (b->blocks_per_cu > 0 ? b->blocks_per_cu
                                : b->staged     ? (b->wpb == 4 ? kBatchStaged4PerCu : kBatchStagedPerCu)
                                                : std::max(1, 32 / b->wpb));
  PSM_HIP(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  hipError_t e = psm::launch_fill_zero(b->counter.as<float>(), psm::kBatchXcds, st);
  for (auto& g : b->groups) {
    if (e != hipSuccess) break;
    e = psm::launch_pack(b->storage, g->data, g->pairs.p, g->scale.as<float>(), g->n_il, g->n_xl,
                         g->n_t, st);
  }
  if (e == hipSuccess)
    e = psm::launch_batch(b->storage, b->staged, b->stage_g, b->d_args.as<psm::MigArgs>(),
                          b->d_segs.as<psm::BatchSeg>(), (int)b->segs.size(),
                          b->counter.as<unsigned>(), b->d_totals.as<unsigned>(), b->wpb, blocks,
                          st);
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(st, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail(c, PSM_ERR_HIP, "batch graph capture failed: %s",
                hipGetErrorString(e != hipSuccess ? e : e2));
  }
  b->graph = g;
  PSM_HIP(c, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return PSM_OK;
}

// Builds the work queue for workgroups of wpb waves (latest t0 tiles of all
// volumes first), uploads it and captures the graph.  The kernel form follows
// the context's staging switch as psm_migrate_volume does: the staged form
// (workgroups of 4 or 8 output crosslines) by default or when forced on, the
// gather form when switched off or when a row of the pair layout exceeds
// kLdsMaxRowBytes.
// stage_g: crossline offsets per staged group (2 or 4); <= 0: the context's
// setting when staging is forced on, else kAutoStageSteps.
int batch_set_shape(psm_batch* b, int wpb, int blocks_per_cu, int stage_g = 0) {
  psm_ctx* c = b->ctx;
  if (wpb != 1 && wpb != 2 && wpb != 4 && wpb != 8)
    return fail(c, PSM_ERR_ARG, "batch waves_per_block must be 1, 2, 4 or 8");
  if (blocks_per_cu < 0 || blocks_per_cu > 64) return fail(c, PSM_ERR_ARG, "bad blocks_per_cu");
  b->staged = c->tune.staging != 0;
  // the staged form needs every item's pair-layout rows within kLdsMaxRowBytes
  for (const psm::MigArgs& a : b->args)
    if ((size_t)(a.n_xd + 1) * a.n_t * (b->storage == PSM_STORE_F32 ? 8 : 4) > psm::kLdsMaxRowBytes)
      b->staged = false;
  if (b->staged) {
    // 4 or 8 output crosslines per workgroup, 2 or 4 offsets per group
    wpb = (wpb == 4 || wpb == 8) ? wpb : psm::kAutoStageWaves;
    if (stage_g != 2 && stage_g != 4)
      stage_g = c->tune.staging > 0 ? c->tune.stage_steps : psm::kAutoStageSteps;
    b->stage_g = stage_g;
  }
  b->wpb = wpb;
  b->blocks_per_cu = blocks_per_cu;
  b->segs.clear();
  int max_tiles = 0;
  for (int t : b->n_tiles) max_tiles = std::max(max_tiles, t);
  // segments in queue order (latest t0 tiles of all volumes first); each
  // segment's items split into kBatchXcds contiguous chunks, chunk x appended
  // to XCD x's queue
  long long first[psm::kBatchXcds] = {};
  for (int r = 0; r < max_tiles; ++r)
    for (int k = 0; k < b->n_items; ++k) {
      if (b->n_tiles[k] <= r) continue;
      const psm::MigArgs& a = b->args[k];
      const int njg = (a.n_out_xl + wpb - 1) / wpb;
      const long long cnt = (long long)a.n_out_il * njg;
      if (cnt > 0x7fffffffLL) return fail(c, PSM_ERR_ARG, "batch queue too long");
      psm::BatchSeg sg{};
      sg.item = k;
      sg.tile = b->n_tiles[k] - 1 - r;
      sg.count = (int)cnt;
      sg.n_jgroups = njg;
      for (int x = 0; x < psm::kBatchXcds; ++x) {
        const long long lo = cnt * x / psm::kBatchXcds, hi = cnt * (x + 1) / psm::kBatchXcds;
        if (first[x] + (hi - lo) > 0x7fffffffLL) return fail(c, PSM_ERR_ARG, "batch queue too long");
        sg.xbase[x] = (int)lo;
        sg.xfirst[x] = (int)first[x];
        first[x] += hi - lo;
      }
      b->segs.push_back(sg);
    }
  for (int x = 0; x < psm::kBatchXcds; ++x) b->totals[x] = (unsigned)first[x];
  hipStream_t st = c->own_stream;
  PSM_HIP(c, upload(b->d_args, b->args.data(), b->args.size() * sizeof(psm::MigArgs), st));
  PSM_HIP(c, upload(b->d_segs, b->segs.data(), b->segs.size() * sizeof(psm::BatchSeg), st));
  PSM_HIP(c, upload(b->d_totals, b->totals, sizeof b->totals, st));
  PSM_HIP(c, hipStreamSynchronize(st));
  return batch_capture(b);
}

int batch_run_locked(psm_batch* b, float* ms) {
  psm_ctx* c = b->ctx;
  PSM_HIP(c, hipSetDevice(c->device));
  hipStream_t st = c->stream;
  if (ms) {
    if (!b->ev0) PSM_HIP(c, hipEventCreate(&b->ev0));
    if (!b->ev1) PSM_HIP(c, hipEventCreate(&b->ev1));
    PSM_HIP(c, hipEventRecord(b->ev0, st));
  }
  PSM_HIP(c, hipGraphLaunch(b->exec, st));
  if (ms) {
    PSM_HIP(c, hipEventRecord(b->ev1, st));
    PSM_HIP(c, hipEventSynchronize(b->ev1));
    PSM_HIP(c, hipEventElapsedTime(ms, b->ev0, b->ev1));
  }
  return PSM_OK;
}

}  // namespace

extern "C" {

int psm_batch_create(psm_ctx* c, const psm_batch_item* items, int n_items, int storage,
                     psm_batch** out) {
  if (!c || !out || !items || n_items <= 0) return c ? fail(c, PSM_ERR_ARG, "bad batch arguments") : PSM_ERR_ARG;
  if (storage != PSM_STORE_F32 && storage != PSM_STORE_F16 && storage != PSM_STORE_BF16)
    return fail(c, PSM_ERR_ARG, "unknown storage %d", storage);
  std::lock_guard<std::mutex> lk(c->mu);
  PSM_HIP(c, hipSetDevice(c->device));
  std::unique_ptr<psm_batch> b(new (std::nothrow) psm_batch());
  if (!b) return fail(c, PSM_ERR_NOMEM, "host allocation failed");
  b->ctx = c;
  b->storage = storage;
  b->n_items = n_items;
  PSM_HIP(c, hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, c->device));
  hipStream_t st = c->own_stream;
  for (int k = 0; k < n_items; ++k) {
    const psm_batch_item& it = items[k];
    const psm_survey* s = &it.survey;
    if (!it.data || !it.vel || !it.out) return fail(c, PSM_ERR_ARG, "batch item %d: null pointer", k);
    if (!(it.alpha > 0)) return fail(c, PSM_ERR_ARG, "batch item %d: alpha must be > 0", k);
    psm_plan plan;
    int rc = plan_impl(c, s, it.vel, PSM_DEVICE, &plan, it.alpha);
    if (rc) return rc;
    struct PlanGuard {
      psm_plan* p;
      ~PlanGuard() { psm_plan_free(p); }
    } guard{&plan};
    if (((size_t)s->n_xl + 1) * s->n_t * sizeof(float2) >= (size_t)INT32_MAX)
      return fail(c, PSM_ERR_ARG, "batch item %d: one inline exceeds 2 GiB", k);
    // data cube group (items migrating the same cube share its packing)
    psm_batch::Group* grp = nullptr;
    for (auto& g : b->groups)
      if (g->data == it.data && g->n_il == (int)s->n_il && g->n_xl == (int)s->n_xl &&
          g->n_t == (int)s->n_t)
        grp = g.get();
    if (!grp) {
      b->groups.emplace_back(new psm_batch::Group{it.data, (int)s->n_il, (int)s->n_xl, (int)s->n_t, {}, {}});
      grp = b->groups.back().get();
      PSM_ALLOC(c, grp->pairs, psm::pairs_bytes(storage, grp->n_il, grp->n_xl, grp->n_t));
      PSM_ALLOC(c, grp->scale, (size_t)(grp->n_il + 1) * (grp->n_xl + 1) * sizeof(float));
    }
    std::vector<psm::Task> tasks(plan.n_out_il);
    for (size_t io = 0; io < plan.n_out_il; ++io)
      tasks[io] = psm::Task{plan.out_to_in_il[io], plan.window[2 * io], plan.window[2 * io + 1],
                            plan.out_to_in_il[io]};
    b->item_bufs.emplace_back(new DevBuf());
    DevBuf& tb = *b->item_bufs.back();
    b->item_bufs.emplace_back(new DevBuf());
    DevBuf& t0b = *b->item_bufs.back();
    PSM_HIP(c, upload(tb, tasks.data(), tasks.size() * sizeof(psm::Task), st));
    PSM_HIP(c, upload(t0b, plan.t0_times, s->n_t * sizeof(float), st));
    psm::MigArgs a{};
    a.xp = grp->pairs.p;
    a.scale = grp->scale.as<float>();
    a.n_rows = (int)s->n_il;
    a.storage = storage;
    a.vel = it.vel;
    a.out = it.out;
    a.t0 = t0b.as<float>();
    a.tasks = tb.as<psm::Task>();
    a.n_t = (int)s->n_t;
    a.n_xd = (int)s->n_xl;
    a.xd_off = s->crossline_padding;
    a.n_xv = (int)s->n_xl;
    a.xv_off = s->crossline_padding;
    a.K = (int)(s->n_xl + 2 * (size_t)s->crossline_padding);
    a.n_out_xl = (int)plan.n_out_xl;
    a.n_out_il = (int)plan.n_out_il;
    const float dtf = (float)s->dt;
    a.inv_dt = 1.0f / dtf;
    a.dx = (float)s->inline_step;
    a.dy = (float)s->crossline_step;
    a.inv_dy = 1.0f / a.dy;
    a.tan2 = plan.tan_theta_sq;
    a.amp = s->amp_correction ? 1 : 0;
    a.d_lo = 0;
    a.d_hi = 1 << 30;
    a.alpha = it.alpha;
    a.alpha_mask = it.alpha;
    b->args.push_back(a);
    b->n_tiles.push_back((int)((s->n_t + 63) / 64));
  }
  PSM_ALLOC(c, b->counter, 256);
  PSM_HIP(c, hipStreamSynchronize(st));
  // the queue and the graph depend on the workgroup width
  const int rc = batch_set_shape(b.get(), 4, 0);
  if (rc) {
    batch_release_graph(b.get());
    return rc;
  }
  *out = b.release();
  return PSM_OK;
}

int psm_batch_set_shape(psm_batch* b, int waves_per_block, int blocks_per_cu) {
  if (!b) return PSM_ERR_ARG;
  std::lock_guard<std::mutex> lk(b->ctx->mu);
  PSM_HIP(b->ctx, hipSetDevice(b->ctx->device));
  return batch_set_shape(b, waves_per_block, blocks_per_cu);
}

int psm_batch_get_shape(psm_batch* b, int* staged, int* waves_per_block, int* blocks_per_cu,
                        int* steps) {
  if (!b) return PSM_ERR_ARG;
  std::lock_guard<std::mutex> lk(b->ctx->mu);
  const int wpb = b->wpb;
  const int bpc = b->blocks_per_cu > 0 ? b->blocks_per_cu
                  : b->staged         ? (wpb == 4 ? kBatchStaged4PerCu : kBatchStagedPerCu)
                                      : std::max(1, 32 / wpb);
  if (staged) *staged = b->staged ? 1 : 0;
  if (waves_per_block) *waves_per_block = wpb;
  if (blocks_per_cu) *blocks_per_cu = bpc;
  if (steps) *steps = b->staged ? b->stage_g : 0;
  return PSM_OK;
}

int psm_batch_run(psm_batch* b, float* ms) {
  if (!b) return PSM_ERR_ARG;
  std::lock_guard<std::mutex> lk(b->ctx->mu);
  return batch_run_locked(b, ms);
}

int psm_batch_autotune(psm_batch* b, int* waves_per_block, int* blocks_per_cu, float* best_ms) {
  if (!b) return PSM_ERR_ARG;
  std::lock_guard<std::mutex> lk(b->ctx->mu);
  psm_ctx* c = b->ctx;
  PSM_HIP(c, hipSetDevice(c->device));
  // candidate launch shapes: workgroup width (output crosslines) x resident
  // workgroups per CU, and for the staged form also the crossline offsets
  // per staged group (the gather form: 0)
  const int cand_gather[][3] = {{1, 32, 0}, {2, 16, 0}, {4, 8, 0}, {8, 4, 0}, {2, 8, 0}, {4, 4, 0}};
  const int cand_staged[][3] = {{8, kBatchStagedPerCu, 4}, {8, kBatchStagedPerCu - 1, 4},
                                {8, kBatchStagedPerCu, 2}, {4, kBatchStaged4PerCu, 4},
                                {4, kBatchStaged4PerCu, 2}, {4, kBatchStaged4PerCu - 1, 4}};
  const bool staged = b->staged;  // the form batch_set_shape chose (it keeps it per shape)
  const int n_cand = 6;
  float best = -1.0f;
  int bw = 4, bb = 0, bg = 0;
  bool warm = false;
  for (int ci = 0; ci < n_cand; ++ci) {
    const int* cd = staged ? cand_staged[ci] : cand_gather[ci];
    int rc = batch_set_shape(b, cd[0], cd[1], cd[2]);
    if (rc) return rc;
    float ms = 0.0f;
    if (!warm) {  // one untimed replay first: code object load, first-touch of the outputs
      rc = batch_run_locked(b, nullptr);
      if (rc) return rc;
      warm = true;
    }
    rc = batch_run_locked(b, &ms);
    if (rc) return rc;
    if (best < 0.0f || ms < best) {
      best = ms;
      bw = cd[0];
      bb = cd[1];
      bg = cd[2];
    }
  }
  const int rc = batch_set_shape(b, bw, bb, bg);
  if (rc) return rc;
  if (waves_per_block) *waves_per_block = bw;
  if (blocks_per_cu) *blocks_per_cu = bb;
  if (best_ms) *best_ms = best;
  return PSM_OK;
}

void psm_batch_destroy(psm_batch* b) {
  if (!b) return;
  psm_ctx* c = b->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  (void)hipStreamSynchronize(c->own_stream);
  batch_release_graph(b);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  for (auto& g : b->groups) {
    g->pairs.release();
    g->scale.release();
  }
  for (auto& d : b->item_bufs) d->release();
  b->d_args.release();
  b->d_segs.release();
  b->d_totals.release();
  b->counter.release();
  delete b;
}

}  // extern "C"
