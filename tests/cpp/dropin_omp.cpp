// dropin_omp.cpp -- a caller of the migration-kernel call shaped like the
// reference driver's output-inline loop (main.cpp:254-383), linked against
// libpsm_hip.so through the C++ shim include/psm_migration_kernel.hpp.
//
// Test program, built by tests/cpp/Makefile (which __graft_entry__.build()
// runs) and run by tests/test_gpu_dropin_cpp.py.  It reads a cube and its Vrms
// cube (raw float32 [n_il][n_xl][n_t], labels 1..n), walks the padded output
// inlines in an OpenMP `parallel for schedule(dynamic, 1)` as main.cpp:255
// does, builds each call's padded aperture window and velocity slice as
// migration_kernel.cpp:385-511 define them, calls
// psm::MigrationKernel::kirchhoffKernel3DWithPadding from every thread, and
// writes the padded output volume.  A failed call is reported as
// "Error: <message>" with exit status 1 (main.cpp:438-441).
//
//   dropin_omp data.bin vel.bin out.bin n_il n_xl n_t dt dx dy pad_il pad_xl
//              angle_deg amp
#include <omp.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "psm_migration_kernel.hpp"

namespace {

std::vector<float> read_all(const char* path, size_t n) {
  std::vector<float> v(n);
  FILE* f = std::fopen(path, "rb");
  if (!f || std::fread(v.data(), sizeof(float), n, f) != n)
    throw std::runtime_error(std::string("cannot read ") + path);
  std::fclose(f);
  return v;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 14) {
    std::fprintf(stderr, "usage: %s data vel out n_il n_xl n_t dt dx dy pad_il pad_xl angle amp\n",
                 argv[0]);
    return 2;
  }
  try {
    const size_t n_il = std::strtoul(argv[4], nullptr, 10);
    const size_t n_xl = std::strtoul(argv[5], nullptr, 10);
    const size_t n_t = std::strtoul(argv[6], nullptr, 10);
    const double dt = std::atof(argv[7]);
    const double dx = std::atof(argv[8]), dy = std::atof(argv[9]);
    const int pad_il = std::atoi(argv[10]), pad_xl = std::atoi(argv[11]);
    const double angle = std::atof(argv[12]);
    const bool amp = std::atoi(argv[13]) != 0;
    const std::vector<float> data = read_all(argv[1], n_il * n_xl * n_t);
    const std::vector<float> vel = read_all(argv[2], n_il * n_xl * n_t);

    // maximum velocity from 9 probes at the last sample, aperture in inlines
    float vmax = 0.0f;
    const size_t pi[3] = {0, n_il / 2, n_il - 1}, px[3] = {0, n_xl / 2, n_xl - 1};
    for (size_t a : pi)
      for (size_t b : px) vmax = std::max(vmax, vel[(a * n_xl + b) * n_t + n_t - 1]);
    if (vmax < 100.0f) vmax = 3000.0f;
    const float tan_theta = (float)std::tan(angle * M_PI / 180.0);
    const float tan_theta_sq = tan_theta * tan_theta;
    const float radius = (vmax * (float)((double)(n_t - 1) * dt) / 2.0f) * std::sqrt(tan_theta_sq);
    const int ap = (int)((double)radius / dx) + 1;
    std::vector<float> t0_times(n_t);
    for (size_t i = 0; i < n_t; ++i) t0_times[i] = (float)((double)i * dt);

    // output inlines with padding (labels 1..n_il extended by pad_il each side)
    const long n_out_il = (long)n_il + 2L * pad_il;
    const size_t K = n_xl + 2 * (size_t)pad_xl;
    std::vector<float> out((size_t)n_out_il * K * n_t, 0.0f);
    std::mutex err_mu;
    std::string err;

#pragma omp parallel for schedule(dynamic, 1)
    for (long io = 0; io < n_out_il; ++io) {
      try {
        // padded output inlines use the nearest real inline
        const long label = io - pad_il + 1;
        const long c = label < 1 ? 0 : (label > (long)n_il ? (long)n_il - 1 : label - 1);
        const long s_buf = std::max(0L, std::max(0L, c - ap) - pad_il);
        const long e_buf = std::min((long)n_il, std::min((long)n_il, c + ap + 1) + pad_il);
        const size_t nb = (size_t)(e_buf - s_buf);
        std::vector<float> buf(nb * K * n_t, 0.0f);
        std::vector<int32_t> il_pos(nb), xl_pos(K);
        for (size_t b = 0; b < nb; ++b) {
          il_pos[b] = (int32_t)(s_buf + (long)b);
          for (size_t x = 0; x < n_xl; ++x)
            std::copy(&data[(((size_t)s_buf + b) * n_xl + x) * n_t],
                      &data[(((size_t)s_buf + b) * n_xl + x) * n_t] + n_t,
                      &buf[(b * K + pad_xl + x) * n_t]);
        }
        for (size_t k = 0; k < K; ++k) xl_pos[k] = (int32_t)k;
        std::vector<float> vslice(K * n_t, 0.0f);
        for (size_t x = 0; x < n_xl; ++x)
          std::copy(&vel[((size_t)c * n_xl + x) * n_t], &vel[((size_t)c * n_xl + x) * n_t] + n_t,
                    &vslice[(pad_xl + x) * n_t]);
        std::vector<std::vector<float>> slice(K, std::vector<float>(n_t, 0.0f));
        psm::MigrationKernel::kirchhoffKernel3DWithPadding(
            buf.data(), nb, K, n_t, vslice.data(), slice, il_pos, xl_pos,
            (int32_t)(c - s_buf), 0, (float)dt, t0_times, tan_theta_sq, (float)dx, (float)dy,
            amp);
        for (size_t j = 0; j < K; ++j)
          std::copy(slice[j].begin(), slice[j].end(), &out[((size_t)io * K + j) * n_t]);
      } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(err_mu);
        if (err.empty()) err = e.what();
      }
    }
    if (!err.empty()) throw std::runtime_error(err);
    FILE* f = std::fopen(argv[3], "wb");
    if (!f || std::fwrite(out.data(), sizeof(float), out.size(), f) != out.size())
      throw std::runtime_error(std::string("cannot write ") + argv[3]);
    std::fclose(f);
    std::printf("migrated %ld output inlines on %d threads\n", n_out_il, omp_get_max_threads());
  } catch (const std::exception& e) {
    std::fprintf(stderr, "Error: %s\n", e.what());
    return 1;
  }
  return 0;
}
