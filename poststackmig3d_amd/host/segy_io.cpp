// segy_io.cpp -- SEG-Y ingest and output of the poststackmig driver.
//
// Ingest (SegyFile, segy_utils.cpp:39-111, 152-214, 243-318): the file is
// mapped once, the trace headers give the label table, and every trace is
// decoded exactly once into a dense [n_il][n_xl][n_t] cube -- the reference
// instead re-reads each inline's traces for every output inline whose aperture
// window contains it (migration_kernel.cpp:441-467 under main.cpp:296-307).
//
// Output (segy_utils.cpp:398-487, 598-648; main.cpp:385-415): one sequential
// pass writing the header block and every output trace in order.
#include <fcntl.h>
#include <omp.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <iostream>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "host_internal.h"

namespace psmh {

namespace {
thread_local std::string g_err;

uint16_t be16(const unsigned char* p) { return (uint16_t)((p[0] << 8) | p[1]); }
int32_t be32(const unsigned char* p) {
  uint32_t v;
  std::memcpy(&v, p, 4);
  return (int32_t)bswap32(v);
}
}  // namespace

void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

// segy_utils.cpp:243-255: value = fraction / 2^24 * 16^(exponent - 64), in
// double, rounded once to float.
float ibm_to_ieee(uint32_t ibm) {
  if (ibm == 0) return 0.0f;
  const bool neg = (ibm >> 31) & 1u;
  const int e = (int)((ibm >> 24) & 0x7Fu) - 64;
  const double m = (double)(ibm & 0x00FFFFFFu) / (double)0x01000000;
  const float v = (float)std::ldexp(m, 4 * e);
  return neg ? -v : v;
}

// segy_utils.cpp:257-270: hexadecimal normalisation by exact scalings of 16,
// then the fraction truncated to 24 bits.  The reference loops forever on an
// infinity and converts a NaN through an undefined float-to-int cast; here
// both encode as the largest IBM magnitude (sign kept for infinities).
uint32_t ieee_to_ibm(float v) {
  if (v == 0.0f) return 0u;
  if (std::isnan(v)) return 0x7FFFFFFFu;
  const uint32_t sign = v < 0.0f ? 0x80000000u : 0u;
  if (std::isinf(v)) return sign | 0x7FFFFFFFu;
  if (v < 0.0f) v = -v;
  int e = 64;
  while (v < 1.0f) {
    v *= 16.0f;
    --e;
  }
  while (v >= 1.0f) {
    v /= 16.0f;
    ++e;
  }
  const uint32_t frac = (uint32_t)(v * 16777216.0f) & 0x00FFFFFFu;
  return sign | ((uint32_t)e << 24) | frac;
}

}  // namespace psmh

using psmh::be16;
using psmh::be32;

int64_t psmh_segy::find(int32_t a, int32_t b) const {
  const auto i = std::lower_bound(il.begin(), il.end(), a);
  const auto j = std::lower_bound(xl.begin(), xl.end(), b);
  if (i == il.end() || *i != a || j == xl.end() || *j != b) return -1;
  return trace_of[(size_t)(i - il.begin()) * xl.size() + (size_t)(j - xl.begin())];
}

namespace {

void open_segy(const char* path, psmh_segy** out) {
  if (!path || !out) throw std::runtime_error("null argument");
  *out = nullptr;
  std::unique_ptr<psmh_segy> s(new psmh_segy);
  s->path = path;
  const int fd = ::open(path, O_RDONLY);
  if (fd < 0) throw std::runtime_error(std::string("Cannot open SEGY file: ") + path);
  struct stat st;
  if (fstat(fd, &st) != 0) {
    ::close(fd);
    throw std::runtime_error(std::string("Cannot open SEGY file: ") + path);
  }
  const size_t len = (size_t)st.st_size;
  if (len < psmh::kText) {
    ::close(fd);
    throw std::runtime_error("Failed to read text header");
  }
  if (len < psmh::kText + psmh::kBinary) {
    ::close(fd);
    throw std::runtime_error("Failed to read binary header");
  }
  void* m = mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fd, 0);
  ::close(fd);
  if (m == MAP_FAILED) throw std::runtime_error(std::string("Cannot map SEGY file: ") + path);
  s->map = static_cast<const unsigned char*>(m);
  s->map_len = len;
  const unsigned char* bh = s->map + psmh::kText;
  s->dt_us = be16(bh + 16);
  if (s->dt_us == 0) {
    psmh_segy_close(s.release());
    throw std::runtime_error("Sample interval (dt) is zero in binary header");
  }
  s->n_t = be16(bh + 20);
  if (s->n_t == 0) {
    psmh_segy_close(s.release());
    throw std::runtime_error("Number of samples per trace is zero in binary header");
  }
  const int fmt = be16(bh + 24);
  if (fmt != 1 && fmt != 5)
    std::cerr << "Warning: Unknown data format code: " << fmt << ", assuming IEEE Float (5)"
              << std::endl;
  s->fmt = (fmt == 1) ? 1 : 5;
  const size_t tr_len = psmh::kTraceHeader + 4 * s->n_t;
  s->n_traces = (len - psmh::kText - psmh::kBinary) / tr_len;
  if (s->n_traces == 0) {
    psmh_segy_close(s.release());
    throw std::runtime_error("No traces found in SEGY file");
  }
  // label table (segy_utils.cpp:163-214): zero labels skipped, later wins
  std::vector<int32_t> ti(s->n_traces), tx(s->n_traces);
  for (size_t k = 0; k < s->n_traces; ++k) {
    const unsigned char* h = s->trace(k);
    ti[k] = be32(h + 188);
    tx[k] = be32(h + 192);
  }
  for (size_t k = 0; k < s->n_traces; ++k)
    if (ti[k] != 0 && tx[k] != 0) {
      s->il.push_back(ti[k]);
      s->xl.push_back(tx[k]);
    }
  std::sort(s->il.begin(), s->il.end());
  s->il.erase(std::unique(s->il.begin(), s->il.end()), s->il.end());
  std::sort(s->xl.begin(), s->xl.end());
  s->xl.erase(std::unique(s->xl.begin(), s->xl.end()), s->xl.end());
  s->trace_of.assign(s->il.size() * s->xl.size(), -1);
  for (size_t k = 0; k < s->n_traces; ++k) {
    if (ti[k] == 0 || tx[k] == 0) continue;
    const size_t a = (size_t)(std::lower_bound(s->il.begin(), s->il.end(), ti[k]) - s->il.begin());
    const size_t b = (size_t)(std::lower_bound(s->xl.begin(), s->xl.end(), tx[k]) - s->xl.begin());
    s->trace_of[a * s->xl.size() + b] = (int64_t)k;
  }
  *out = s.release();
}

void read_cube(psmh_segy* s, float* cube) {
  if (!s || !cube) throw std::runtime_error("null argument");
  const size_t n_il = s->il.size(), n_xl = s->xl.size(), n_t = s->n_t;
  const bool ibm = s->fmt == 1;
#pragma omp parallel for schedule(static)
  for (long long r = 0; r < (long long)(n_il * n_xl); ++r) {
    float* dst = cube + (size_t)r * n_t;
    const int64_t k = s->trace_of[(size_t)r];
    if (k < 0) {
      std::memset(dst, 0, n_t * sizeof(float));
      continue;
    }
    const unsigned char* src = s->trace((size_t)k) + psmh::kTraceHeader;
    for (size_t i = 0; i < n_t; ++i) {
      uint32_t w;
      std::memcpy(&w, src + 4 * i, 4);
      w = psmh::bswap32(w);
      if (ibm) {
        dst[i] = psmh::ibm_to_ieee(w);
      } else {
        float f;
        std::memcpy(&f, &w, 4);
        dst[i] = f;
      }
    }
  }
}

void write_segy(const char* path, const psmh_segy* in, const int32_t* oil, size_t n_oil,
                const int32_t* oxl, size_t n_oxl, const float* image, size_t n_t) {
  if (!path || !in || !oil || !oxl || !image) throw std::runtime_error("null argument");
  if (n_t != in->n_t) throw std::runtime_error("image sample count differs from the input file");
  if (n_t > 65535) throw std::runtime_error("too many samples per trace for SEG-Y");
  FILE* f = std::fopen(path, "wb");
  if (!f) throw std::runtime_error(std::string("Cannot create output SEGY file: ") + path);
  std::unique_ptr<FILE, int (*)(FILE*)> guard(f, &std::fclose);
  // header block: text copied, binary copied with samples @20 and format @24
  std::vector<unsigned char> head(in->map, in->map + psmh::kText + psmh::kBinary);
  unsigned char* bh = head.data() + psmh::kText;
  bh[20] = (unsigned char)(n_t >> 8);
  bh[21] = (unsigned char)(n_t & 0xFF);
  bh[24] = 0;
  bh[25] = (unsigned char)in->fmt;
  if (std::fwrite(head.data(), 1, head.size(), f) != head.size())
    throw std::runtime_error("write failed: " + std::string(path));
  const size_t tr_len = psmh::kTraceHeader + 4 * n_t;
  // Output inlines in chunks of about 16 MB: a chunk is encoded (OpenMP over
  // its traces) into one of two buffers while a second thread writes the
  // previous chunk from the other; the file is written in inline order.
  const size_t line_bytes = n_oxl * tr_len;
  const size_t chunk_lines = std::max<size_t>(1, ((size_t)16 << 20) / std::max<size_t>(line_bytes, 1));
  std::vector<unsigned char> bufs[2];
  for (auto& b : bufs) b.resize(std::min(chunk_lines, n_oil) * line_bytes);
  std::thread writer;
  bool write_ok = true;
  auto join_writer = [&] {
    if (writer.joinable()) writer.join();
  };
  struct Joiner {  // (joins on every exit, so no thread outlives the buffers)
    std::function<void()> fn;
    ~Joiner() { fn(); }
  } joiner{join_writer};
  for (size_t c0 = 0, slot = 0; c0 < n_oil; c0 += chunk_lines, slot ^= 1) {
    const size_t nl = std::min(chunk_lines, n_oil - c0);
    unsigned char* const chunk = bufs[slot].data();
#pragma omp parallel for schedule(static)
    for (long long q = 0; q < (long long)(nl * n_oxl); ++q) {
      const size_t io = c0 + (size_t)q / n_oxl;
      const size_t jo = (size_t)q % n_oxl;
      const int32_t il_label = oil[io];
      const bool real_il = std::binary_search(in->il.begin(), in->il.end(), il_label);
      unsigned char* t = chunk + (size_t)q * tr_len;
      const int64_t k = real_il ? in->find(il_label, oxl[jo]) : -1;
      if (k >= 0) {
        std::memcpy(t, in->trace((size_t)k), psmh::kTraceHeader);
      } else {
        std::memset(t, 0, psmh::kTraceHeader);
        const uint32_t a = psmh::bswap32((uint32_t)il_label), b = psmh::bswap32((uint32_t)oxl[jo]);
        std::memcpy(t + 188, &a, 4);
        std::memcpy(t + 192, &b, 4);
      }
      const float* src = image + (io * n_oxl + (size_t)jo) * n_t;
      unsigned char* d = t + psmh::kTraceHeader;
      for (size_t i = 0; i < n_t; ++i) {
        uint32_t w;
        if (in->fmt == 1) {
          w = psmh::ieee_to_ibm(src[i]);
        } else {
          std::memcpy(&w, src + i, 4);
        }
        w = psmh::bswap32(w);
        std::memcpy(d + 4 * i, &w, 4);
      }
    }
    join_writer();  // (the previous chunk is written; its buffer is the next one's)
    if (!write_ok) throw std::runtime_error("write failed: " + std::string(path));
    writer = std::thread([&write_ok, f, chunk, n = nl * line_bytes] {
      if (std::fwrite(chunk, 1, n, f) != n) write_ok = false;
    });
  }
  join_writer();
  if (!write_ok) throw std::runtime_error("write failed: " + std::string(path));
  if (std::fflush(f) != 0) throw std::runtime_error("write failed: " + std::string(path));
}

}  // namespace

extern "C" {

const char* psmh_last_error(void) { return psmh::last_error(); }

int psmh_segy_open(const char* path, psmh_segy** out) {
  return psmh::guarded([&] { open_segy(path, out); });
}

void psmh_segy_close(psmh_segy* s) {
  if (!s) return;
  if (s->map) munmap(const_cast<unsigned char*>(s->map), s->map_len);
  delete s;
}

int psmh_segy_info(const psmh_segy* s, size_t* n_traces, size_t* n_t, int* dt_us, int* fmt,
                   size_t* n_il, size_t* n_xl) {
  if (!s) return PSM_ERR_ARG;
  if (n_traces) *n_traces = s->n_traces;
  if (n_t) *n_t = s->n_t;
  if (dt_us) *dt_us = s->dt_us;
  if (fmt) *fmt = s->fmt;
  if (n_il) *n_il = s->il.size();
  if (n_xl) *n_xl = s->xl.size();
  return PSM_OK;
}

const int32_t* psmh_segy_inline_labels(const psmh_segy* s) { return s ? s->il.data() : nullptr; }
const int32_t* psmh_segy_crossline_labels(const psmh_segy* s) { return s ? s->xl.data() : nullptr; }

int psmh_segy_read_cube(psmh_segy* s, float* cube) {
  return psmh::guarded([&] { read_cube(s, cube); });
}

void psmh_ibm_to_ieee(const uint32_t* ibm, float* out, size_t n) {
  for (size_t i = 0; i < n; ++i) out[i] = psmh::ibm_to_ieee(ibm[i]);
}

void psmh_ieee_to_ibm(const float* x, uint32_t* out, size_t n) {
  for (size_t i = 0; i < n; ++i) out[i] = psmh::ieee_to_ibm(x[i]);
}

int psmh_segy_write(const char* path, const psmh_segy* in, const int32_t* out_il_labels,
                    size_t n_out_il, const int32_t* out_xl_labels, size_t n_out_xl,
                    const float* image, size_t n_t) {
  return psmh::guarded([&] {
    write_segy(path, in, out_il_labels, n_out_il, out_xl_labels, n_out_xl, image, n_t);
  });
}

}  // extern "C"
