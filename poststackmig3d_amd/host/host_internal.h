// host_internal.h -- shared pieces of libpsm_host.so.
#ifndef PSM_HOST_INTERNAL_H
#define PSM_HOST_INTERNAL_H

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "psm_host.h"

namespace psmh {

// Thread-local message of the last failure (psmh_last_error).
void set_error(const std::string& msg);

// Runs f(), mapping exceptions to a status code and the error message.
template <class F>
int guarded(F&& f) {
  try {
    f();
    set_error("");
    return PSM_OK;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return PSM_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_error(e.what());
    return PSM_ERR_ARG;
  }
}

inline uint32_t bswap32(uint32_t v) { return __builtin_bswap32(v); }
inline uint16_t bswap16(uint16_t v) { return __builtin_bswap16(v); }

float ibm_to_ieee(uint32_t ibm);
uint32_t ieee_to_ibm(float v);

constexpr size_t kText = 3200, kBinary = 400, kTraceHeader = 240;

}  // namespace psmh

// An opened SEG-Y file: the whole file mapped read-only, its metadata and the
// (inline, crossline) -> trace table as a dense index over the sorted labels.
struct psmh_segy {
  std::string path;
  const unsigned char* map = nullptr;
  size_t map_len = 0;
  size_t n_traces = 0, n_t = 0;
  int dt_us = 0, fmt = 5;
  std::vector<int32_t> il, xl;        // sorted distinct labels
  std::vector<int64_t> trace_of;      // [n_il * n_xl] file trace index, -1 = missing
  const unsigned char* trace(size_t k) const {
    return map + psmh::kText + psmh::kBinary + k * (psmh::kTraceHeader + 4 * n_t);
  }
  // file trace index of labels (il, xl), -1 if absent
  int64_t find(int32_t il_label, int32_t xl_label) const;
};

#endif  // PSM_HOST_INTERNAL_H
