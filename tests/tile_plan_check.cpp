// CPU check of the LDS-tiled force kernel's index arithmetic
// (tfm_amd/csrc/tile_plan.hpp and the block/plane mapping of k_force):
// for bead positions read from a file it builds both grids exactly as the
// binning does, replays every block, and verifies that
//   - every bead slot belongs to exactly one block,
//   - window cells and slots stay inside cstart / spos,
//   - every neighbour row a bead scans lies inside its plane's window, and
//     every LDS index stays inside the staged range and the LDS budget.
// Usage: tile_plan_check <positions.f32> <N> <fine csl> <coarse csl> <d_m> <het mask file>
#include "../tfm_amd/csrc/tile_plan.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

using namespace ccpmi;

struct Geom {
  float csl;
  int cps, half, n_c;
};

static Geom geom(float csl, float d_m) {
  Geom g;
  g.csl = csl;
  g.cps = 2 * (int)std::ceil(d_m / csl);
  if (g.cps < 4) g.cps = 4;
  g.half = g.cps / 2;
  g.n_c = g.cps * g.cps * g.cps;
  return g;
}

static int coord(const Geom &g, float x) {
  int c = (int)std::floor(x / g.csl) + g.half;
  return c < 0 ? 0 : (c >= g.cps ? g.cps - 1 : c);
}

static int fails = 0;
#define CHECK(c)                                                                                                       \
  do {                                                                                                                 \
    if (!(c)) {                                                                                                        \
      if (fails++ < 20) std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c);                             \
    }                                                                                                                  \
  } while (0)

struct Built {
  std::vector<uint32_t> cstart, order; // order[s] = bead in slot s
};

static Built build(const Geom &g, const std::vector<float> &r, const std::vector<int> &members) {
  Built b;
  std::vector<uint32_t> count(g.n_c, 0);
  std::vector<int> key(members.size());
  for (size_t k = 0; k < members.size(); ++k) {
    const int i = members[k];
    key[k] = (coord(g, r[3 * i + 2]) * g.cps + coord(g, r[3 * i + 1])) * g.cps + coord(g, r[3 * i]);
    ++count[key[k]];
  }
  b.cstart.assign(g.n_c + 1, 0);
  for (int c = 0; c < g.n_c; ++c) b.cstart[c + 1] = b.cstart[c] + count[c];
  std::vector<uint32_t> fill(b.cstart.begin(), b.cstart.end() - 1);
  b.order.assign(members.size(), 0);
  for (size_t k = 0; k < members.size(); ++k) b.order[fill[key[k]]++] = (uint32_t)members[k];
  return b;
}

int main(int argc, char **argv) {
  if (argc < 7) return 2;
  const int N = std::atoi(argv[2]);
  const float csl_f = (float)std::atof(argv[3]), csl_h = (float)std::atof(argv[4]), d_m = (float)std::atof(argv[5]);
  std::vector<float> r(3 * (size_t)N);
  std::ifstream(argv[1], std::ios::binary).read(reinterpret_cast<char *>(r.data()), 12 * (size_t)N);
  std::vector<char> het(N);
  std::ifstream(argv[6], std::ios::binary).read(het.data(), N);
  const Geom gf = geom(csl_f, d_m), gh = geom(csl_h, d_m);
  std::vector<int> all(N), hets;
  for (int i = 0; i < N; ++i) {
    all[i] = i;
    if (het[i]) hets.push_back(i);
  }
  const Built F = build(gf, r, all), H = build(gh, r, hets);
  const int cps2 = gf.cps * gf.cps;
  const int blocks = (N + TILE - 1) / TILE + gf.cps;
  std::vector<int> owner(N, 0);
  long lds_blocks = 0, fallback_f = 0, fallback_h = 0;
  for (int b = 0; b < blocks; ++b) {
    // block -> plane: planes round their bead count up to whole tiles
    uint32_t acc = 0;
    int plane = -1;
    uint32_t beg = 0, end = 0;
    for (int z = 0; z < gf.cps; ++z) {
      const uint32_t pb = F.cstart[z * cps2], pe = F.cstart[(z + 1) * cps2];
      const uint32_t nb = (pe - pb + TILE - 1) / TILE;
      if ((uint32_t)b >= acc && (uint32_t)b < acc + nb) {
        plane = z;
        beg = pb + ((uint32_t)b - acc) * TILE;
        end = std::min(beg + (uint32_t)TILE, pe);
        break;
      }
      acc += nb;
    }
    if (plane < 0) continue;
    CHECK(beg < end && end <= (uint32_t)N);
    for (uint32_t s = beg; s < end; ++s) ++owner[s];
    const int ya = coord(gf, r[3 * F.order[beg] + 1]), yb = coord(gf, r[3 * F.order[end - 1] + 1]);
    // fine windows
    Window wf[3];
    int pos_off = 0, cell_off = 0;
    for (int k = 0; k < 3; ++k) {
      Window W = make_window(gf.cps, plane - 1 + k, ya, yb);
      CHECK(W.c0 >= 0 && W.c1 <= gf.n_c && W.c0 <= W.c1);
      W.s0 = W.plane >= 0 ? F.cstart[W.c0] : 0u;
      W.s1 = W.plane >= 0 ? F.cstart[W.c1] : 0u;
      W.pos_base = pos_off;
      W.cell_base = cell_off;
      if (W.plane >= 0) {
        pos_off += (int)(W.s1 - W.s0);
        cell_off += W.c1 - W.c0 + 1;
      }
      wf[k] = W;
    }
    const bool fok = pos_off <= LDS_POS && cell_off <= LDS_CELLS;
    if (!fok) pos_off = cell_off = 0, ++fallback_f;
    // het windows from the het beads' coarse bounding box
    int box[4] = {1 << 30, -1, 1 << 30, -1};
    for (uint32_t s = beg; s < end; ++s) {
      const int i = F.order[s];
      if (!het[i]) continue;
      const int y = coord(gh, r[3 * i + 1]), z = coord(gh, r[3 * i + 2]);
      box[0] = std::min(box[0], y), box[1] = std::max(box[1], y), box[2] = std::min(box[2], z),
      box[3] = std::max(box[3], z);
    }
    Window wh[MAX_WIN];
    bool hok = false;
    if (box[1] >= 0 && box[3] - box[2] + 3 <= MAX_WIN) {
      hok = true;
      for (int k = 0; k < box[3] - box[2] + 3; ++k) {
        Window W = make_window(gh.cps, box[2] - 1 + k, box[0], box[1]);
        CHECK(W.c0 >= 0 && W.c1 <= gh.n_c && W.c0 <= W.c1);
        W.s0 = W.plane >= 0 ? H.cstart[W.c0] : 0u;
        W.s1 = W.plane >= 0 ? H.cstart[W.c1] : 0u;
        W.pos_base = pos_off;
        W.cell_base = cell_off;
        if (W.plane >= 0) {
          pos_off += (int)(W.s1 - W.s0);
          cell_off += W.c1 - W.c0 + 1;
        }
        wh[k] = W;
      }
      if (pos_off > LDS_POS || cell_off > LDS_CELLS) hok = false;
    }
    if (box[1] >= 0 && !hok) ++fallback_h;
    if (fok || hok) ++lds_blocks;
    // every bead's rows stay inside its windows
    auto rows_inside = [&](const Geom &g, const Built &B, const Window *win, int zfirst, int nwin, int cx, int cy,
                           int cz, int pos_limit, int cell_limit) {
      for (int rr = 0; rr < 9; ++rr) {
        const int yy = cy + rr % 3 - 1, zz = cz + rr / 3 - 1;
        int rc0, rc1;
        row_cells(g.cps, cx, yy, zz, rc0, rc1);
        if (rc1 <= rc0) continue;
        const int k = zz - zfirst;
        CHECK(k >= 0 && k < nwin);
        if (k < 0 || k >= nwin) continue;
        const Window &W = win[k];
        CHECK(W.plane == zz);
        CHECK(rc0 >= W.c0 && rc1 <= W.c1);
        const int ci0 = W.cell_base + (rc0 - W.c0), ci1 = W.cell_base + (rc1 - W.c0);
        CHECK(ci0 >= 0 && ci1 < cell_limit && ci1 <= W.cell_base + (W.c1 - W.c0));
        const uint32_t lo = B.cstart[rc0], hi = B.cstart[rc1];
        CHECK(lo >= W.s0 && hi <= W.s1);
        if (hi > lo) {
          const int p0 = W.pos_base + (int)(lo - W.s0), p1 = W.pos_base + (int)(hi - 1 - W.s0);
          CHECK(p0 >= 0 && p1 < pos_limit);
        }
      }
    };
    for (uint32_t s = beg; s < end; ++s) {
      const int i = F.order[s];
      const int cx = coord(gf, r[3 * i]), cy = coord(gf, r[3 * i + 1]), cz = coord(gf, r[3 * i + 2]);
      CHECK(cz == plane);
      CHECK(cy >= ya && cy <= yb);
      if (fok) rows_inside(gf, F, wf, plane - 1, 3, cx, cy, cz, LDS_POS, LDS_CELLS);
      if (het[i] && hok) {
        const int hx = coord(gh, r[3 * i]), hy = coord(gh, r[3 * i + 1]), hz = coord(gh, r[3 * i + 2]);
        rows_inside(gh, H, wh, box[2] - 1, box[3] - box[2] + 3, hx, hy, hz, LDS_POS, LDS_CELLS);
      }
    }
  }
  for (int s = 0; s < N; ++s) CHECK(owner[s] == 1);
  std::printf("blocks=%d lds_blocks=%ld fine_fallback=%ld het_fallback=%ld fails=%d\n", blocks, lds_blocks,
              fallback_f, fallback_h, fails);
  return fails ? 1 : 0;
}
