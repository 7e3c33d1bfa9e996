// ccp-analyze: drop-in for ccp's analysis executable
// (Program/src/ccp-analyze.cu:8-136) on top of the libccpmi C ABI.
//
//   ccp-analyze <sim_dir>
//
// Reads adjustable-parameters.dat, then for every simulation i_s the
// initial condition initial-condition-XXX.gro and the trajectory files
// trajectory-XXX-YYY.trr; writes analysis-XXX.dat per simulation, then
// analysis-fin.dat and contact-map.bin. A log "<unix time>.log" is written in
// the working directory and removed on success.

#include "ccpmi.h"

#include <glob.h>

#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>

namespace {

struct Logger {
  std::fstream file;
  bool to_file = false;

  void open(const std::string &path) {
    file.open(path, std::ios::out | std::ios::app);
    file.close();
    file.open(path, std::ios::in | std::ios::out | std::ios::ate);
    to_file = file.is_open();
  }
  void close() {
    if (file.is_open()) file.close();
    to_file = false;
  }
  void record(const std::string &msg) {
    time_t now = time(nullptr);
    char ts[21];
    strftime(ts, 21, "[%d/%m/%y %H:%M:%S] ", localtime(&now));
    if (to_file) file << ts << msg << std::endl;
    std::cout << ts << msg << std::endl;
  }
  void progress(const std::string &line) {
    if (to_file) {
      file << line;
      file.seekp(-(std::streamoff)line.size(), std::ios::cur);
    }
    std::cout << line << "\r";
    std::cout.flush();
  }
};

void log_cb(const char *line, void *user) {
  Logger *lg = static_cast<Logger *>(user);
  std::string s(line);
  if (s.rfind("progress: ", 0) == 0) lg->progress(s);
  else lg->record(s);
}

std::string pad3(unsigned v) {
  std::ostringstream s;
  s << std::setw(3) << std::setfill('0') << v;
  return s.str();
}

unsigned glob_count(const std::string &pat) {
  glob_t g;
  int rv = glob(pat.c_str(), 0, nullptr, &g);
  unsigned n = 0;
  if (rv == 0) n = (unsigned)g.gl_pathc;
  globfree(&g);
  if (rv != 0 && rv != GLOB_NOMATCH) throw std::runtime_error("unable to find matches of " + pat);
  return n;
}

void check(ccp_sim *h, int rc) {
  if (rc != CCP_OK) throw std::runtime_error(ccp_last_error(h));
}

void check(ccp_analysis *a, int rc) {
  if (rc != CCP_OK) throw std::runtime_error(ccp_analysis_last_error(a));
}

} // namespace

int main(int argc, const char **argv) {
  if (argc < 2) {
    std::cout << "no arguments\n";
    return EXIT_FAILURE;
  }
  if (argc > 2) {
    std::cout << "extra arguments\n";
    return EXIT_FAILURE;
  }
  const std::string sim_dir = argv[1];
  const time_t t_s = time(nullptr);
  const std::string log_path = std::to_string(t_s) + ".log";
  Logger lg;
  lg.open(log_path);
  ccp_sim *h = nullptr;
  ccp_analysis *a = nullptr;
  try {
    std::string path = sim_dir + "/adjustable-parameters.dat";
    std::ifstream in(path);
    if (!in.is_open()) throw std::runtime_error("unable to open " + path);
    std::stringstream text;
    text << in.rdbuf();
    in.close();
    h = ccp_create_logged(text.str().c_str(), log_cb, &lg);
    if (!h) throw std::runtime_error(ccp_last_error(nullptr));
    check(h, ccp_analysis_create(h, &a));

    const unsigned n_s = glob_count(sim_dir + "/initial-condition-*");
    for (unsigned i_s = 0; i_s < n_s; ++i_s) {
      check(h, ccp_read_frame_txt(h, (sim_dir + "/initial-condition-" + pad3(i_s) + ".gro").c_str()));
      check(a, ccp_analysis_add_frame(a));
      const unsigned n_t_f = glob_count(sim_dir + "/trajectory-" + pad3(i_s) + "*");
      for (unsigned i_t_f = 0; i_t_f < n_t_f; ++i_t_f) {
        const std::string trr = sim_dir + "/trajectory-" + pad3(i_s) + "-" + pad3(i_t_f) + ".trr";
        check(a, ccp_analysis_add_trajectory_file(a, trr.c_str()));
        lg.record("added trajectory file " + pad3(i_s) + "-" + pad3(i_t_f));
      }
      check(a, ccp_analysis_end_simulation(a, (sim_dir + "/analysis-" + pad3(i_s) + ".dat").c_str()));
    }
    check(a, ccp_analysis_finish(a, (sim_dir + "/analysis-fin.dat").c_str(), (sim_dir + "/contact-map.bin").c_str()));
  } catch (const std::exception &e) {
    lg.record(e.what());
    ccp_analysis_destroy(a);
    ccp_destroy(h);
    return EXIT_FAILURE;
  }
  ccp_analysis_destroy(a);
  ccp_destroy(h);
  lg.close();
  std::remove(log_path.c_str());
  return EXIT_SUCCESS;
}
