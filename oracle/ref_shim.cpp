// ref_shim.cpp -- C entry points around the reference's own compiled kernel.
//
// TEST INFRASTRUCTURE ONLY (the checker / CPU baseline).  This file is ours;
// it is compiled together with the reference's unmodified sources
// (src/migration_kernel.cpp and the files it links against) by
// oracle/Makefile into oracle/_ref/libpsm_ref.so.  Nothing in the product
// links it.
//
//   psmr_kernel_inline   calls MigrationKernel::kirchhoffKernel3DWithPadding
//                        (migration_kernel.h:16-31) once, with the nested
//                        OpenMP team forced to one thread as it is when
//                        main.cpp:255 calls it from its parallel loop.
//   psmr_migrate_volume  main.cpp:254-383's loop over output inlines
//                        (schedule(dynamic,1)), each iteration building the
//                        padded aperture window of migration_kernel.cpp:385-468
//                        and the velocity slice of :470-511 from an in-memory
//                        cube instead of a SEG-Y file, then calling the
//                        reference kernel.  This is the kernel-only CPU path.
#include <omp.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "migration_kernel.h"

extern "C" {

__attribute__((visibility("default"))) int psmr_kernel_inline(
    const float* in, size_t n_buf_il, size_t n_buf_xl, size_t n_t,
    const float* vel, float* out, size_t n_xl, const int32_t* il_pos,
    size_t n_il_pos, int32_t cur_il_pos, int32_t cur_xl_start, float dt,
    const float* t0_times, float tan_theta_sq, float dx, float dy, int amp) {
  std::vector<std::vector<float>> slice(n_xl, std::vector<float>(n_t, 0.0f));
  std::vector<int32_t> ilp(il_pos, il_pos + n_il_pos);
  std::vector<int32_t> xlp(n_buf_xl);
  for (size_t i = 0; i < n_buf_xl; ++i) xlp[i] = (int32_t)i;
  std::vector<float> t0(t0_times, t0_times + n_t);
  const int saved = omp_get_max_threads();
  omp_set_num_threads(1);
  MigrationKernel::kirchhoffKernel3DWithPadding(
      in, n_buf_il, n_buf_xl, n_t, vel, slice, ilp, xlp, cur_il_pos,
      cur_xl_start, dt, t0, tan_theta_sq, dx, dy, amp != 0);
  omp_set_num_threads(saved);
  for (size_t j = 0; j < n_xl; ++j)
    std::memcpy(out + j * n_t, slice[j].data(), n_t * sizeof(float));
  return 0;
}

// data, vel: [n_il][n_xl][n_t]; out: [n_out_il][n_out_xl][n_t];
// c_of[io]: index of the real inline output inline io maps to (main.cpp:258-288);
// rows (optional): subset of output inlines to compute.
__attribute__((visibility("default"))) int psmr_migrate_volume(
    const float* data, const float* vel, size_t n_il, size_t n_xl, size_t n_t,
    double dt, double il_step, double xl_step, int pad_il, int pad_xl,
    float tan_theta_sq, int max_ap_il, int amp, const int32_t* c_of,
    float* out, size_t n_out_il, size_t n_out_xl, const int32_t* rows,
    size_t n_rows, int nthreads) {
  std::vector<float> t0(n_t);
  for (size_t i = 0; i < n_t; ++i) t0[i] = static_cast<float>(i * dt);
  const size_t K = n_xl + 2 * static_cast<size_t>(pad_xl);
  const size_t nr = rows ? n_rows : n_out_il;
  std::vector<int32_t> xlp(K);
  for (size_t i = 0; i < K; ++i) xlp[i] = (int32_t)i;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  omp_set_max_active_levels(1);
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t r = 0; r < nr; ++r) {
    const size_t io = rows ? static_cast<size_t>(rows[r]) : r;
    const int c = c_of[io];
    const int s_req = std::max(0, c - max_ap_il);
    const int e_req = std::min(static_cast<int>(n_il), c + max_ap_il + 1);
    const int s_buf = std::max(0, s_req - pad_il);
    const int e_buf = std::min(static_cast<int>(n_il), e_req + pad_il);
    const size_t nb = static_cast<size_t>(e_buf - s_buf);
    std::vector<float> buf(nb * K * n_t, 0.0f);
    std::vector<int32_t> ilp(nb);
    for (int b = s_buf; b < e_buf; ++b) {
      ilp[b - s_buf] = b;
      for (size_t x = 0; x < n_xl; ++x)
        std::memcpy(&buf[((size_t)(b - s_buf) * K + pad_xl + x) * n_t],
                    data + ((size_t)b * n_xl + x) * n_t, n_t * sizeof(float));
    }
    std::vector<float> vs(K * n_t, 0.0f);
    for (size_t x = 0; x < n_xl; ++x)
      std::memcpy(&vs[(pad_xl + x) * n_t], vel + ((size_t)c * n_xl + x) * n_t,
                  n_t * sizeof(float));
    int32_t cur = -1;
    for (size_t i = 0; i < nb; ++i)
      if (ilp[i] == c) { cur = (int32_t)i; break; }
    if (cur < 0) cur = 0;
    std::vector<std::vector<float>> slice(n_out_xl, std::vector<float>(n_t, 0.0f));
    MigrationKernel::kirchhoffKernel3DWithPadding(
        buf.data(), nb, K, n_t, vs.data(), slice, ilp, xlp, cur, 0,
        static_cast<float>(dt), t0, tan_theta_sq, static_cast<float>(il_step),
        static_cast<float>(xl_step), amp != 0);
    for (size_t j = 0; j < n_out_xl; ++j)
      std::memcpy(out + (io * n_out_xl + j) * n_t, slice[j].data(), n_t * sizeof(float));
  }
  return 0;
}

// Bounded CPU sample: for each listed output inline, the reference kernel is
// called once for the output crosslines [xl0, xl0 + m) only
// (current_xl_start = xl0, an output slice of m crosslines); one call per
// OpenMP thread as in main.cpp:255.  out: [n_rows][m][n_t].
__attribute__((visibility("default"))) int psmr_migrate_sample(
    const float* data, const float* vel, size_t n_il, size_t n_xl, size_t n_t,
    double dt, double il_step, double xl_step, int pad_il, int pad_xl,
    float tan_theta_sq, int max_ap_il, int amp, const int32_t* c_of,
    const int32_t* rows, size_t n_rows, int xl0, size_t m, float* out, int nthreads) {
  std::vector<float> t0(n_t);
  for (size_t i = 0; i < n_t; ++i) t0[i] = static_cast<float>(i * dt);
  const size_t K = n_xl + 2 * static_cast<size_t>(pad_xl);
  std::vector<int32_t> xlp(K);
  for (size_t i = 0; i < K; ++i) xlp[i] = (int32_t)i;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  omp_set_max_active_levels(1);
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t r = 0; r < n_rows; ++r) {
    const int c = c_of[rows[r]];
    const int s_req = std::max(0, c - max_ap_il);
    const int e_req = std::min(static_cast<int>(n_il), c + max_ap_il + 1);
    const int s_buf = std::max(0, s_req - pad_il);
    const int e_buf = std::min(static_cast<int>(n_il), e_req + pad_il);
    const size_t nb = static_cast<size_t>(e_buf - s_buf);
    std::vector<float> buf(nb * K * n_t, 0.0f);
    std::vector<int32_t> ilp(nb);
    for (int b = s_buf; b < e_buf; ++b) {
      ilp[b - s_buf] = b;
      for (size_t x = 0; x < n_xl; ++x)
        std::memcpy(&buf[((size_t)(b - s_buf) * K + pad_xl + x) * n_t],
                    data + ((size_t)b * n_xl + x) * n_t, n_t * sizeof(float));
    }
    std::vector<float> vs(K * n_t, 0.0f);
    for (size_t x = 0; x < n_xl; ++x)
      std::memcpy(&vs[(pad_xl + x) * n_t], vel + ((size_t)c * n_xl + x) * n_t,
                  n_t * sizeof(float));
    const int32_t cur = (int32_t)(c - s_buf);
    std::vector<std::vector<float>> slice(m, std::vector<float>(n_t, 0.0f));
    MigrationKernel::kirchhoffKernel3DWithPadding(
        buf.data(), nb, K, n_t, vs.data(), slice, ilp, xlp, cur, xl0,
        static_cast<float>(dt), t0, tan_theta_sq, static_cast<float>(il_step),
        static_cast<float>(xl_step), amp != 0);
    for (size_t j = 0; j < m; ++j)
      std::memcpy(out + (r * m + j) * n_t, slice[j].data(), n_t * sizeof(float));
  }
  return 0;
}

}  // extern "C"
