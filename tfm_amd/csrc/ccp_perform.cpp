// ccp-perform: drop-in for ccp's simulation executable
// (Program/src/ccp-perform.cu:8-128) on top of the libccpmi C ABI.
//
//   ccp-perform <sim_dir>        new simulation: initial condition, one
//                                trajectory file of frames_per_file frames,
//                                checkpoint
//   ccp-perform <sim_dir> <i_s>  continue simulation i_s from its checkpoint
//
// Files in <sim_dir>: adjustable-parameters.dat (input),
// initial-condition-XXX.gro, lamina-binding-sites-XXX.gro,
// trajectory-XXX-YYY.trr, checkpoint-XXX.bin. A log "<unix time>.log" is
// written in the working directory and removed on success.

#include "ccpmi.h"

#include <glob.h>

#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
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
  static std::string stamp() {
    time_t now = time(nullptr);
    char ts[21];
    strftime(ts, 21, "[%d/%m/%y %H:%M:%S] ", localtime(&now));
    return ts;
  }
  void record(const std::string &msg) {
    std::string ts = stamp();
    if (to_file) file << ts << msg << std::endl;
    std::cout << ts << msg << std::endl;
  }
  // "progress: 050.0%" overwritten in place (util.cu:39-50)
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

struct Fail : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void check(ccp_sim *h, int rc) {
  if (rc != CCP_OK) throw Fail(ccp_last_error(h));
}

} // namespace

int main(int argc, const char **argv) {
  if (argc < 2) {
    std::cout << "no arguments\n";
    return EXIT_FAILURE;
  }
  if (argc > 3) {
    std::cout << "extra arguments\n";
    return EXIT_FAILURE;
  }
  const std::string sim_dir = argv[1];
  const bool new_sim = argc == 2;
  const time_t t_s = time(nullptr);
  const std::string log_path = std::to_string(t_s) + ".log";
  Logger lg;
  lg.open(log_path);
  ccp_sim *h = nullptr;
  try {
    std::string path = sim_dir + "/adjustable-parameters.dat";
    std::ifstream in(path);
    if (!in.is_open()) throw Fail("unable to open " + path);
    std::stringstream text;
    text << in.rdbuf();
    in.close();
    h = ccp_create_logged(text.str().c_str(), log_cb, &lg);
    if (!h) throw Fail(ccp_last_error(nullptr));

    unsigned i_s, i_t_f;
    if (new_sim) {
      i_s = glob_count(sim_dir + "/initial-condition-*");
      i_t_f = 0;
      check(h, ccp_generate_initial_condition(h));
      check(h, ccp_write_frame_txt(h, (sim_dir + "/initial-condition-" + pad3(i_s) + ".gro").c_str()));
      check(h, ccp_write_lbs_txt(h, (sim_dir + "/lamina-binding-sites-" + pad3(i_s) + ".gro").c_str()));
    } else {
      i_s = (unsigned)std::stoi(argv[2]);
      i_t_f = glob_count(sim_dir + "/trajectory-" + pad3(i_s) + "*");
      check(h, ccp_load_checkpoint(h, (sim_dir + "/checkpoint-" + pad3(i_s) + ".bin").c_str()));
    }
    lg.record("i_s = " + pad3(i_s) + " i_t_f = " + pad3(i_t_f) + " ");
    check(h, ccp_run_simulation(h, (sim_dir + "/trajectory-" + pad3(i_s) + "-" + pad3(i_t_f) + ".trr").c_str()));
    check(h, ccp_save_checkpoint(h, (sim_dir + "/checkpoint-" + pad3(i_s) + ".bin").c_str()));
  } catch (const std::exception &e) {
    lg.record(e.what());
    ccp_destroy(h);
    return EXIT_FAILURE;
  }
  ccp_destroy(h);
  lg.close();
  std::remove(log_path.c_str());
  return EXIT_SUCCESS;
}
