// Writes Philox4x32-10 vectors computed by rocRAND's engine (host pass of
// the installed header), an implementation independent of this project's.
// Output: one line per vector, "c0 c1 c2 c3 k0 k1 o0 o1 o2 o3" in hex.
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdint>

// ten_rounds is protected in the engine; a derived type re-exports it.
struct open_engine : rocrand_device::philox4x32_10_engine {
  using rocrand_device::philox4x32_10_engine::ten_rounds;
};

int main() {
  open_engine eng;
  uint32_t s = 0x12345678u;
  auto next = [&]() { s ^= s << 13; s ^= s >> 17; s ^= s << 5; return s; };
  const uint32_t fixed[][6] = {
      {0, 0, 0, 0, 0, 0},
      {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
      {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u},
      {1, 0, 0, 0, 3, 0},
      {131071, 199999, 0, 1, 3, 0},
  };
  for (auto &v : fixed) {
    uint4 c = {v[0], v[1], v[2], v[3]};
    uint2 k = {v[4], v[5]};
    uint4 o = eng.ten_rounds(c, k);
    printf("%08x %08x %08x %08x %08x %08x %08x %08x %08x %08x\n", v[0], v[1], v[2], v[3], v[4], v[5], o.x, o.y, o.z, o.w);
  }
  for (int i = 0; i < 64; ++i) {
    uint32_t v[6];
    for (auto &x : v) x = next();
    uint4 c = {v[0], v[1], v[2], v[3]};
    uint2 k = {v[4], v[5]};
    uint4 o = eng.ten_rounds(c, k);
    printf("%08x %08x %08x %08x %08x %08x %08x %08x %08x %08x\n", v[0], v[1], v[2], v[3], v[4], v[5], o.x, o.y, o.z, o.w);
  }
  return 0;
}
