// ccp-profile: drop-in for ccp's timing executable
// (Program/src/ccp-profile.cu:8-105) on top of the libccpmi C ABI.
//
//   ccp-profile <out_file>
//
// Eight runs of ccp's own force field with lamina binding sites, N = 2^8 ..
// 2^15 beads at chromatin volume fraction 0.125 and LBS area fraction 0.25,
// ccp's default timestep (1/2048) and frames of 1/dt steps; run i has
// 128 - 16 i frames. Each run generates its initial condition, then times
// the whole simulation call (every frame copied to the host and written as a
// .trr frame to /dev/null). Writes one line per run, "NNNNN t.ttt": N and
// the wall-clock milliseconds per step, as ccp's profiler.

#include "ccpmi.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>

namespace {

// ccp's zero-padded number formatting (util.cu cnfs): width `len`, `prec`
// decimals, padded on the left with `fill`
std::string cnfs(double v, int len, char fill, int prec) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.*f", prec, v);
  std::string s(buf);
  while ((int)s.size() < len) s.insert(s.begin(), fill);
  return s;
}

void quiet(const char *, void *) {}

} // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::cout << "no arguments\n";
    return EXIT_FAILURE;
  }
  if (argc > 2) {
    std::cout << "extra arguments\n";
    return EXIT_FAILURE;
  }
  const std::string path = argv[1];
  std::ofstream out(path);
  if (!out.is_open()) {
    std::cout << "unable to open " << path << "\n";
    return EXIT_FAILURE;
  }
  const unsigned n_t = 8;
  const double cvf = 0.125, laf = 0.25;
  const double rco = 1.154701, lco = 0.5, dt = 1.0 / 2048.0;
  for (unsigned i_t = 0; i_t < n_t; ++i_t) {
    const unsigned N = 1u << (8 + i_t);
    const float R_n = (float)(rco / 2.0 + rco / 2.0 * std::pow(N / cvf, 1.0 / 3));
    const unsigned n_l = (unsigned)(laf * 4.0 / std::pow(lco / (R_n - rco), 2.0));
    const unsigned fpf = 128 - 16 * i_t;
    std::ostringstream par;
    par << "number_of_particles " << cnfs(N, 5, '0', 0) << "\n";
    par << "nucleus_radius " << cnfs(R_n, 5, '0', 2) << "\n";
    par << "number_of_lbs " << cnfs(n_l, 5, '0', 0) << "\n";
    par << "frames_per_file " << cnfs(fpf, 4, '0', 0) << "\n";
    ccp_sim *h = ccp_create_logged(par.str().c_str(), quiet, nullptr);
    if (!h) {
      std::cout << ccp_last_error(nullptr) << "\n";
      return EXIT_FAILURE;
    }
    if (ccp_generate_initial_condition(h) != CCP_OK) {
      std::cout << ccp_last_error(h) << "\n";
      ccp_destroy(h);
      return EXIT_FAILURE;
    }
    const auto t0 = std::chrono::system_clock::now();
    const int rc = ccp_run_simulation(h, "/dev/null");
    const auto t1 = std::chrono::system_clock::now();
    if (rc != CCP_OK) {
      std::cout << ccp_last_error(h) << "\n";
      ccp_destroy(h);
      return EXIT_FAILURE;
    }
    ccp_destroy(h);
    const double t_e = 1000.0 * std::chrono::duration<double>(t1 - t0).count() / (fpf / dt);
    out << cnfs(N, 5, '0', 0) << " " << cnfs(t_e, 5, '0', 3) << std::endl;
    std::cout << cnfs(N, 5, '0', 0) << " " << cnfs(t_e, 5, '0', 3) << std::endl;
  }
  return EXIT_SUCCESS;
}
