// cli.cpp -- the `poststackmig <config>` run (main.cpp:66-443) over the C ABI:
// configuration, one-pass ingest of the input cube into page-locked memory,
// the Vrms cube from the velocity provider, the whole padded volume migrated
// on the GPU by psm_migrate_volume (which computes the output labels, the
// nearest-inline mapping, the aperture window and the sticky parameters of
// main.cpp:140-217, 254-383 itself), and the output SEG-Y.
#include <omp.h>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <memory>
#include <string>
#include <thread>

#include "host_internal.h"

namespace {

struct HostBuf {
  void* p = nullptr;
  ~HostBuf() { psm_free_host(p); }
  float* alloc(size_t n) {
    if (psm_alloc_host(n * sizeof(float), &p) != PSM_OK)
      throw std::runtime_error("page-locked host allocation of " + std::to_string(n * sizeof(float)) +
                               " bytes failed");
    return static_cast<float*>(p);
  }
};

struct Ctx {
  psm_ctx* c = nullptr;
  ~Ctx() { psm_destroy(c); }
};

struct Segy {
  psmh_segy* s = nullptr;
  ~Segy() { psmh_segy_close(s); }
};

struct Plan {
  psm_plan p{};
  bool owned = false;
  ~Plan() {
    if (owned) psm_plan_free(&p);
  }
};

void check(int rc, const std::string& what) {
  if (rc != PSM_OK) throw std::runtime_error(what);
}

void run(const char* config_path, int device) {
  const auto t_start = std::chrono::steady_clock::now();
  const std::string bar(60, '=');
  std::cout << bar << "\nPost-stack 3D Kirchhoff Migration (MI355X)\n" << bar << std::endl;

  psmh_config cfg;
  if (psmh_config_parse(config_path, &cfg) != PSM_OK) throw std::runtime_error(psmh_last_error());
  std::cout << "Input file: " << cfg.input_data << "\nOutput file: " << cfg.output_data
            << "\nInline step: " << cfg.inline_step << " m\nCrossline step: " << cfg.crossline_step
            << " m\nInline padding: " << cfg.inline_padding
            << "\nCrossline padding: " << cfg.crossline_padding
            << "\nAngle aperture: " << cfg.angle_aperture
            << " deg\nAmplitude correction: " << (cfg.amp_correction ? "true" : "false") << std::endl;
  if (cfg.n_threads > 0) omp_set_num_threads(cfg.n_threads);
  std::cout << "Host threads: " << omp_get_max_threads() << ", GPU device: " << device << "\n"
            << bar << std::endl;

  std::cout << "\n[1/5] Reading the input cube..." << std::endl;
  // the device context (HIP start-up) is created while the input file's
  // headers are scanned; it is waited for before the first page-locked
  // buffer is allocated
  Ctx ctx;
  int ctx_rc = PSM_OK;
  std::thread ctx_thread([&] { ctx_rc = psm_create(&ctx.c, device); });
  struct JoinCtx {
    std::thread& t;
    ~JoinCtx() {
      if (t.joinable()) t.join();
    }
  } join_ctx{ctx_thread};
  Segy in;
  if (psmh_segy_open(cfg.input_data, &in.s) != PSM_OK) throw std::runtime_error(psmh_last_error());
  size_t n_traces = 0, n_t = 0, n_il = 0, n_xl = 0;
  int dt_us = 0, fmt = 0;
  psmh_segy_info(in.s, &n_traces, &n_t, &dt_us, &fmt, &n_il, &n_xl);
  if (n_il == 0 || n_xl == 0) throw std::runtime_error("no trace with nonzero inline and crossline labels");
  const double dt = dt_us * 1e-6;
  std::cout << "Lookup table: " << n_traces << " traces, " << n_il << " inlines, " << n_xl
            << " crosslines, data format " << fmt << (fmt == 1 ? " (IBM Float)" : " (IEEE Float)")
            << "\nInput geometry: IL=" << n_il << ", XL=" << n_xl << ", T=" << n_t << ", dt=" << dt << "s"
            << std::endl;
  ctx_thread.join();
  check(ctx_rc, "cannot create a context on GPU " + std::to_string(device));
  HostBuf cube_buf, vel_buf, out_buf;
  float* cube = cube_buf.alloc(n_il * n_xl * n_t);
  if (psmh_segy_read_cube(in.s, cube) != PSM_OK) throw std::runtime_error(psmh_last_error());

  std::cout << "\n[2/5] Building the velocity cube..." << std::endl;
  float* vel = vel_buf.alloc(n_il * n_xl * n_t);
  if (psmh_velocity_cube(ctx.c, cfg.velocity, psmh_segy_inline_labels(in.s), n_il,
                         psmh_segy_crossline_labels(in.s), n_xl, n_t, dt, vel) != PSM_OK)
    throw std::runtime_error(psmh_last_error());

  std::cout << "\n[3/5] Output geometry and migration parameters..." << std::endl;
  psm_survey s{};
  s.n_il = n_il;
  s.n_xl = n_xl;
  s.n_t = n_t;
  s.dt = dt;
  s.inline_step = cfg.inline_step;
  s.crossline_step = cfg.crossline_step;
  s.inline_padding = cfg.inline_padding;
  s.crossline_padding = cfg.crossline_padding;
  s.angle_aperture = cfg.angle_aperture;
  s.amp_correction = cfg.amp_correction;
  s.inline_labels = psmh_segy_inline_labels(in.s);
  s.crossline_labels = psmh_segy_crossline_labels(in.s);
  s.max_velocity = 0.0f;  // estimateMaxVelocity from the Vrms cube (main.cpp:22-64)
  Plan plan;
  const int prc = psm_survey_plan(&s, vel, PSM_HOST, &plan.p);
  if (prc != PSM_OK) throw std::runtime_error("invalid survey geometry (status " + std::to_string(prc) + ")");
  plan.owned = true;
  s.max_velocity = plan.p.max_velocity;
  const double max_t = (double)(n_t - 1) * dt;
  const float radius = (plan.p.max_velocity * (float)max_t / 2.0f) * std::sqrt(plan.p.tan_theta_sq);
  std::cout << "Output dimensions: IL=" << plan.p.n_out_il << ", XL=" << plan.p.n_out_xl << std::fixed
            << std::setprecision(0) << "\nMaximum velocity: " << plan.p.max_velocity << " m/s"
            << std::setprecision(2) << "\nMaximum time: " << max_t << " s" << std::setprecision(0)
            << "\nAperture radius: IL=" << radius << "m, XL=" << radius << "m"
            << "\nBuffer size: IL=+-" << plan.p.max_aperture_il << " lines" << std::endl;
  std::cout.unsetf(std::ios_base::fixed);
  std::cout << std::setprecision(6);

  std::cout << "\n[4/5] Migrating " << plan.p.n_out_il << " output inlines on the GPU..." << std::endl;
  const size_t n_out = plan.p.n_out_il * plan.p.n_out_xl * n_t;
  float* out = out_buf.alloc(n_out);
  const auto t_mig = std::chrono::steady_clock::now();
  if (psm_migrate_volume(ctx.c, &s, cube, vel, out, PSM_HOST) != PSM_OK)
    throw std::runtime_error(psm_last_error(ctx.c));
  const double mig_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_mig).count();
  std::cout << "Migration: " << mig_s << " s (including transfers)" << std::endl;

  std::cout << "\n[5/5] Writing " << cfg.output_data << "..." << std::endl;
  if (psmh_segy_write(cfg.output_data, in.s, plan.p.out_inline_labels, plan.p.n_out_il,
                      plan.p.out_crossline_labels, plan.p.n_out_xl, out, n_t) != PSM_OK)
    throw std::runtime_error(psmh_last_error());

  const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  std::cout << "\nMigration complete!\nTotal time: " << total << " seconds\nProcessed: " << n_il
            << " inlines x " << n_xl << " crosslines x " << n_t << " samples" << std::endl;
}

}  // namespace

extern "C" int psmh_run(const char* config_path, int device) {
  try {
    if (!config_path) throw std::runtime_error("no configuration file");
    run(config_path, device);
    return 0;
  } catch (const std::exception& e) {
    std::cerr << "Error: " << e.what() << std::endl;  // main.cpp:438-441
    return 1;
  }
}
