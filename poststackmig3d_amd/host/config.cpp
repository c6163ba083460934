// config.cpp -- the poststackmig configuration file
// (ConfigParser, config_parser.cpp:10-212; Config, config_parser.h:7-20).
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "host_internal.h"

namespace {

std::string trim(const std::string& s) {
  const size_t a = s.find_first_not_of(" \t\n\r");
  if (a == std::string::npos) return "";
  const size_t b = s.find_last_not_of(" \t\n\r");
  return s.substr(a, b - a + 1);
}

std::string unquote(const std::string& s0) {
  std::string s = trim(s0);
  if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') || (s.front() == '\'' && s.back() == '\'')))
    return s.substr(1, s.size() - 2);
  return s;
}

bool to_bool(const std::string& v) {
  std::string l = v;
  std::transform(l.begin(), l.end(), l.begin(), [](unsigned char c) { return (char)std::tolower(c); });
  l = trim(l);
  if (l == "true" || l == "1" || l == "yes" || l == "on") return true;
  if (l == "false" || l == "0" || l == "no" || l == "off") return false;
  throw std::runtime_error("Cannot parse boolean value: " + v);
}

// std::stod / std::stoi semantics: leading number, rest ignored
double to_double(const std::string& v) {
  try {
    return std::stod(v);
  } catch (const std::exception&) {
    throw std::runtime_error("Cannot parse double value: " + v);
  }
}

int to_int(const std::string& v) {
  try {
    return std::stoi(v);
  } catch (const std::exception&) {
    throw std::runtime_error("Cannot parse int value: " + v);
  }
}

void copy_field(char (&dst)[1024], const std::string& v, const char* name) {
  if (v.size() >= sizeof dst) throw std::runtime_error(std::string(name) + " is too long");
  std::memcpy(dst, v.c_str(), v.size() + 1);
}

void parse(const char* path, psmh_config* c) {
  if (!path || !c) throw std::runtime_error("null argument");
  std::ifstream f(path);
  if (!f.is_open()) throw std::runtime_error(std::string("Config file not found: ") + path);
  std::map<std::string, std::string> kv;
  std::string line;
  int n = 0;
  while (std::getline(f, line)) {
    ++n;
    line = trim(line);
    if (line.empty() || line[0] == '#') continue;
    const size_t eq = line.find('=');
    if (eq == std::string::npos)
      throw std::runtime_error("Invalid format at line " + std::to_string(n) +
                               ": expected key=value, got: " + line);
    kv[trim(line.substr(0, eq))] = unquote(trim(line.substr(eq + 1)));
  }
  std::vector<std::string> missing;
  for (const char* k : {"input_data", "output_data", "inline_step", "crossline_step", "velocity"})
    if (!kv.count(k)) missing.push_back(k);
  if (!missing.empty()) {
    std::string m = "Missing required parameters: ";
    for (size_t i = 0; i < missing.size(); ++i) m += (i ? ", " : "") + missing[i];
    throw std::runtime_error(m);
  }
  std::memset(c, 0, sizeof *c);
  c->angle_aperture = 30.0;
  c->amp_correction = 1;
  try {
    copy_field(c->input_data, kv["input_data"], "input_data");
    copy_field(c->output_data, kv["output_data"], "output_data");
    copy_field(c->velocity, kv["velocity"], "velocity");
    c->inline_step = to_double(kv["inline_step"]);
    c->crossline_step = to_double(kv["crossline_step"]);
    if (kv.count("inline_padding")) c->inline_padding = to_int(kv["inline_padding"]);
    if (kv.count("crossline_padding")) c->crossline_padding = to_int(kv["crossline_padding"]);
    if (kv.count("angle_aperture")) c->angle_aperture = to_double(kv["angle_aperture"]);
    if (kv.count("amp_correction")) c->amp_correction = to_bool(kv["amp_correction"]) ? 1 : 0;
    if (kv.count("n_threads")) c->n_threads = to_int(kv["n_threads"]);
  } catch (const std::exception& e) {
    throw std::runtime_error(std::string("Error parsing config values: ") + e.what());
  }
  // validateConfig (config_parser.cpp:176-212)
  auto num = [](double v) { return std::to_string(v); };
  if (c->inline_step <= 0) throw std::runtime_error("inline_step must be positive, got: " + num(c->inline_step));
  if (c->crossline_step <= 0)
    throw std::runtime_error("crossline_step must be positive, got: " + num(c->crossline_step));
  if (c->inline_padding < 0)
    throw std::runtime_error("inline_padding must be non-negative, got: " + std::to_string(c->inline_padding));
  if (c->crossline_padding < 0)
    throw std::runtime_error("crossline_padding must be non-negative, got: " +
                             std::to_string(c->crossline_padding));
  if (c->angle_aperture <= 0 || c->angle_aperture >= 90)
    throw std::runtime_error("angle_aperture must be in range (0, 90), got: " + num(c->angle_aperture));
  if (c->n_threads < 0)
    throw std::runtime_error("n_threads must be non-negative (0 = use all available), got: " +
                             std::to_string(c->n_threads));
  // The reference's check of a numeric velocity (config_parser.cpp:200-211)
  // throws inside its own try block and is swallowed by its catch, so a
  // non-positive constant is only rejected by the constant velocity provider.
}

}  // namespace

extern "C" int psmh_config_parse(const char* path, psmh_config* out) {
  return psmh::guarded([&] { parse(path, out); });
}
