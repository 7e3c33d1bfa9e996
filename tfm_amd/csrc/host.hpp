// Host utilities shared by the simulation, analysis and batch handles of
// libccpmi: ccp's constants, parameter map and number formatting, errors,
// and a device buffer.
#pragma once

#include "ccpmi.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <iomanip>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>

namespace ccpmi {
namespace host {

// ---------------------------------------------------------------- constants
// ccp model constants (chrdat.cuh:15-24, chrsim.cuh:18-20)
inline constexpr float K_B = 0.003356f;
inline constexpr float L_0 = 1.0f;
inline constexpr float K_E = 128.0f;
inline constexpr float K_BEND = 1.0f;
inline constexpr float E_P = 0.5f;
inline constexpr float RCO = 1.154701f;
inline constexpr float ACO = 2.0f;
inline constexpr float MDL = 21.10139f;
inline constexpr float E_L = 8.0f;
inline constexpr float LCO = 0.5f;
inline constexpr float DT_DEF = 1.0f / 2048;
inline constexpr float MIS = 0.838732f;
inline constexpr uint32_t N_DEF = 30386;
inline const uint32_t CHR_LIM[13] = {0, 2284, 4599, 6687, 9338, 12508, 15193, 17477, 19792, 21880, 24531, 27701, 30386};

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e) {
  if (e != hipSuccess) throw Error(CCP_ERR_HIP, std::string("hip: ") + hipGetErrorString(e));
}

// the message ccp_last_error(NULL) returns on this thread (calls with no handle)
void set_null_error(const std::string &msg);

// whitespace "key value" pairs; the first occurrence of a key wins (util.cu:72-80)
struct ParMap : std::map<std::string, std::string> {
  explicit ParMap(const char *text) {
    std::istringstream is(text ? text : "");
    std::string k, v;
    while (is >> k >> v) insert({k, v});
  }
  template <class T> T get(const std::string &key, T def) const {
    auto it = find(key);
    if (it == end()) return def;
    T val{};
    std::stringstream{it->second} >> val;
    return val;
  }
  bool has(const std::string &k) const { return find(k) != end(); }
};

// cnfs (util.cuh:91-104): formatted number
template <class T> std::string cnfs(T num, int len = 0, char fillc = ' ', int prc = 0) {
  std::stringstream s;
  if (len > 0) s << std::setw(len);
  if (fillc != ' ') s << std::setfill(fillc);
  if (prc > 0) s << std::setprecision(prc) << std::fixed;
  s << num;
  return s.str();
}

template <class T> struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    n = count;
    hip_check(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
  }
  void free_() {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
};


} // namespace host
} // namespace ccpmi
