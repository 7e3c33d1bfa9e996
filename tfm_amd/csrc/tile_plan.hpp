// Index arithmetic of the LDS-tiled force kernel, shared by the kernel and
// a host-side check (tests/tile_plan_check.cpp) that verifies on the CPU
// that every access stays inside its buffers and every neighbour row a
// bead needs lies inside the staged windows.
//
// A block handles up to TILE consecutive beads of one z-plane of the fine
// grid (so its cells span a few y-rows of that plane). For a grid and a set
// of planes, a window is the contiguous run of cells of rows
// [y_lo - 1, y_hi + 1] (clamped) of one plane; cells of a plane are
// contiguous in the x-fastest index, so a window is one contiguous slot range
// of the ordered positions.
#pragma once

#include <stdint.h>

#ifndef __HIPCC__
#define TP_HD inline
#else
#define TP_HD __host__ __device__ __forceinline__
#endif

namespace ccpmi {

constexpr int TILE = 256;       // beads per block
constexpr int MAX_WIN = 4;      // windows per grid (planes z-1 .. z+2)
constexpr int LDS_POS = 3072;   // staged positions per block (float4)
constexpr int LDS_CELLS = 4096; // staged cell starts per block

struct Window {
  int plane;        // z of the window's plane (or -1: empty)
  int c0, c1;       // cells [c0, c1) of the window (c1 - c0 = rows * cps)
  uint32_t s0, s1;  // slots [s0, s1) = [cstart[c0], cstart[c1])
  int pos_base;     // LDS offset of slot s0
  int cell_base;    // LDS offset of cstart[c0] (cells c0..c1 inclusive)
};

// Window of plane zz over rows [ylo-1, yhi+1] (clamped), or an empty one when
// the plane is outside the grid.
TP_HD Window make_window(int cps, int zz, int ylo, int yhi) {
  Window w{};
  if (zz < 0 || zz >= cps) {
    w.plane = -1;
    w.c0 = w.c1 = 0;
    return w;
  }
  const int y0 = ylo - 1 < 0 ? 0 : ylo - 1;
  const int y1 = yhi + 1 > cps - 1 ? cps - 1 : yhi + 1;
  w.plane = zz;
  w.c0 = (zz * cps + y0) * cps;
  w.c1 = (zz * cps + y1 + 1) * cps;
  return w;
}

// Cells [rc0, rc1) of neighbour row (yy, zz) around x = cx, or rc0 = rc1 = 0
// when the row is outside the grid (same clamping as the global-memory scan).
TP_HD void row_cells(int cps, int cx, int yy, int zz, int &rc0, int &rc1) {
  const bool ok = zz >= 0 && zz < cps && yy >= 0 && yy < cps;
  const int row = (zz * cps + yy) * cps;
  rc0 = ok ? row + (cx - 1 < 0 ? 0 : cx - 1) : 0;
  rc1 = ok ? row + (cx + 1 > cps - 1 ? cps - 1 : cx + 1) + 1 : 0;
}

} // namespace ccpmi
