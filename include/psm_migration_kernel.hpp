// psm_migration_kernel.hpp -- header-only C++ shim that gives the reference's
// migration-kernel call, unchanged, on top of the C ABI of libpsm_hip.so.
//
// Replaces: MigrationKernel::kirchhoffKernel3DWithPadding
//           (include/migration_kernel.h:16-31, src/migration_kernel.cpp:17-383)
//
// Same parameter list and meaning; the output vector<vector<float>> is
// pre-sized by the caller ([n_xl][n_t], main.cpp:325-327) and fully
// overwritten.  The call is synchronous and re-entrant: each host thread uses
// its own psm_ctx (one HIP stream + device buffers) on device
// PSM_DEVICE_INDEX (environment, default 0), so main.cpp's
// `#pragma omp parallel for` over output inlines (main.cpp:255) keeps working.
// Errors surface as std::runtime_error (main.cpp:438-441 turns them into
// "Error: ..." and exit status 1).
#ifndef PSM_MIGRATION_KERNEL_HPP
#define PSM_MIGRATION_KERNEL_HPP

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "psm.h"

namespace psm {

namespace detail {

struct CtxDeleter {
  void operator()(psm_ctx* c) const { psm_destroy(c); }
};

inline psm_ctx* thread_context() {
  thread_local std::unique_ptr<psm_ctx, CtxDeleter> ctx;
  if (!ctx) {
    const char* env = std::getenv("PSM_DEVICE_INDEX");
    const int dev = env ? std::atoi(env) : 0;
    psm_ctx* c = nullptr;
    const int rc = psm_create(&c, dev);
    if (rc != PSM_OK)
      throw std::runtime_error("psm_create(device " + std::to_string(dev) +
                               ") failed with code " + std::to_string(rc));
    ctx.reset(c);
  }
  return ctx.get();
}

}  // namespace detail

class MigrationKernel {
 public:
  static void kirchhoffKernel3DWithPadding(
      const float* input_buffer, size_t n_buf_il, size_t n_buf_xl, size_t n_t,
      const float* velocity_slice, std::vector<std::vector<float>>& output_slice,
      const std::vector<int32_t>& il_pos_indices, const std::vector<int32_t>& xl_pos_indices,
      int32_t current_il_pos, int32_t current_xl_start, float dt,
      const std::vector<float>& t0_times, float tan_theta_sq, float inline_bin_size,
      float crossline_bin_size, bool amp_correction) {
    const size_t n_xl = output_slice.size();
    if (n_xl == 0 || n_t == 0) return;  // migration_kernel.cpp:35-37
    if (t0_times.size() < n_t) throw std::runtime_error("t0_times shorter than n_t");
    thread_local std::vector<float> flat;
    flat.resize(n_xl * n_t);
    psm_ctx* c = detail::thread_context();
    const int rc = psm_kirchhoff_inline(
        c, input_buffer, n_buf_il, n_buf_xl, n_t, velocity_slice, flat.data(), n_xl,
        il_pos_indices.data(), il_pos_indices.size(), xl_pos_indices.data(),
        xl_pos_indices.size(), current_il_pos, current_xl_start, dt, t0_times.data(),
        tan_theta_sq, inline_bin_size, crossline_bin_size, amp_correction ? 1 : 0);
    if (rc != PSM_OK) throw std::runtime_error(psm_last_error(c));
    for (size_t j = 0; j < n_xl; ++j) {
      output_slice[j].resize(n_t);
      std::copy(flat.begin() + j * n_t, flat.begin() + (j + 1) * n_t, output_slice[j].begin());
    }
  }
};

}  // namespace psm

#endif  // PSM_MIGRATION_KERNEL_HPP
