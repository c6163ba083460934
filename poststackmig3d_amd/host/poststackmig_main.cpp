// poststackmig_main.cpp -- `poststackmig <config_file>`, the reference's
// command line (main.cpp:66-70) with the migration on the GPU
// (PSM_DEVICE_INDEX selects the device, default 0).
#include <cstdlib>
#include <iostream>

#include "psm_host.h"

int main(int argc, char** argv) {
  if (argc != 2) {
    std::cerr << "Usage: " << argv[0] << " <config_file>" << std::endl;
    return 1;
  }
  const char* env = std::getenv("PSM_DEVICE_INDEX");
  return psmh_run(argv[1], env ? std::atoi(env) : 0);
}
