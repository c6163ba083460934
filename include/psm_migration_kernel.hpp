// psm_migration_kernel.hpp -- header-only C++ shim that gives the reference's
// migration-kernel call, unchanged, on top of the C ABI of libpsm_hip.so.
//
// Replaces: MigrationKernel::kirchhoffKernel3DWithPadding
//           (include/migration_kernel.h:16-31, src/migration_kernel.cpp:17-383)
//
// Same parameter list and meaning; the output vector<vector<float>> is
// pre-sized by the caller ([n_xl][n_t], main.cpp:325-327) and fully
// overwritten.  The call is synchronous and re-entrant, so main.cpp's
// `#pragma omp parallel for` over output inlines (main.cpp:255) keeps working:
// each call leases a psm_ctx (one HIP stream + device buffers) from a
// process-wide pool on device PSM_DEVICE_INDEX (environment, default 0).  The
// pool holds at most PSM_MAX_CONTEXTS contexts (environment, default 4);
// further threads wait for a free one.  Each context keeps the device buffers
// of the largest window it has served (at the 401 x 401 x 1001 survey's
// windows, 375 x 461 x 1001 samples, about 2.1 GB), so the cap bounds device
// memory independently of the OpenMP thread count.  Errors surface as
// std::runtime_error (main.cpp:438-441 turns them into "Error: ..." and exit
// status 1).
#ifndef PSM_MIGRATION_KERNEL_HPP
#define PSM_MIGRATION_KERNEL_HPP

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "psm.h"

namespace psm {

namespace detail {

// Contexts created on demand up to a cap and reused; never destroyed (the
// pool lives until the process ends, so no context outlives the HIP runtime's
// teardown order).
class ContextPool {
 public:
  ContextPool() {
    const char* dev = std::getenv("PSM_DEVICE_INDEX");
    device_ = dev ? std::atoi(dev) : 0;
    const char* cap = std::getenv("PSM_MAX_CONTEXTS");
    cap_ = cap ? std::atoi(cap) : 4;
    if (cap_ < 1) cap_ = 1;
  }

  psm_ctx* acquire() {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [this] { return !idle_.empty() || created_ < cap_; });
    if (!idle_.empty()) {
      psm_ctx* c = idle_.back();
      idle_.pop_back();
      return c;
    }
    ++created_;
    lk.unlock();
    psm_ctx* c = nullptr;
    const int rc = psm_create(&c, device_);
    if (rc != PSM_OK) {
      lk.lock();
      --created_;
      cv_.notify_one();
      throw std::runtime_error("psm_create(device " + std::to_string(device_) +
                               ") failed with code " + std::to_string(rc));
    }
    return c;
  }

  void release(psm_ctx* c) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      idle_.push_back(c);
    }
    cv_.notify_one();
  }

 private:
  std::mutex mu_;
  std::condition_variable cv_;
  std::vector<psm_ctx*> idle_;
  int created_ = 0;
  int cap_ = 4;
  int device_ = 0;
};

inline ContextPool& pool() {
  static ContextPool* p = new ContextPool();
  return *p;
}

// One leased context for the duration of a call.
class Lease {
 public:
  Lease() : c_(pool().acquire()) {}
  ~Lease() { pool().release(c_); }
  Lease(const Lease&) = delete;
  Lease& operator=(const Lease&) = delete;
  psm_ctx* get() const { return c_; }

 private:
  psm_ctx* c_;
};

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
    {
      detail::Lease lease;
      const int rc = psm_kirchhoff_inline(
          lease.get(), input_buffer, n_buf_il, n_buf_xl, n_t, velocity_slice, flat.data(), n_xl,
          il_pos_indices.data(), il_pos_indices.size(), xl_pos_indices.data(),
          xl_pos_indices.size(), current_il_pos, current_xl_start, dt, t0_times.data(),
          tan_theta_sq, inline_bin_size, crossline_bin_size, amp_correction ? 1 : 0);
      // (the message is copied while the context is still leased)
      if (rc != PSM_OK) throw std::runtime_error(psm_last_error(lease.get()));
    }
    for (size_t j = 0; j < n_xl; ++j) {
      output_slice[j].resize(n_t);
      std::copy(flat.begin() + j * n_t, flat.begin() + (j + 1) * n_t, output_slice[j].begin());
    }
  }
};

}  // namespace psm

#endif  // PSM_MIGRATION_KERNEL_HPP
