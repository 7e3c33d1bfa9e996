// gfx950 kernels of the ccp Brownian-dynamics step and of ccp's analysis.
//
// One stochastic-Heun stage is two launches:
//   k_force_bal<MODE, STAGE>  eight lanes per bead, beads in cell order (het
//                         beads first within each workgroup): the bead's
//                         neighbour rows (ccp's 27-cell neighbourhood,
//                         calc_srf, chrsim.cu:277-317, or 125 half-size
//                         cells) trimmed to its own cutoff, their slots
//                         compacted into an LDS list that the eight lanes
//                         share evenly, the partial forces summed by DPP.
//                         Lane 0 then adds the bonded + wall force of the
//                         order pass, runs the Heun update and bins the new
//                         position into the next grid (atomic count -> bucket
//                         slot, aggregated per-scan-block totals).
//   k_force<MODE, STAGE>  the other kernel shapes: one lane per bead, or one
//                         lane per neighbour plane with shuffles summing the
//                         parts; optionally (fine_cell_size < cell_size) short
//                         pairs from a fine all-bead grid and the long het-het
//                         tail from a coarse grid of het beads only.
//   k_order               one thread per cell: exclusive scan of the counts
//                         (block prefix from the per-block totals), then a
//                         rank-by-bead-index ordering of each cell's bucket
//                         entries (positions with the bead index in .w), so
//                         the ordered positions equal a stable sort by cell
//                         (the result of ccp's cub::DeviceRadixSort::SortPairs
//                         over identity values, sugrid.cu:125-133) bit for
//                         bit. Its first workgroups also compute the next
//                         force pass's per-bead terms at one lane per bead:
//                         the next step's normals and the bonded + wall forces.
// Workgroup indices are remapped so each XCD works on one contiguous run of
// the grid (its L2 holds that run's neighbour rows).
// Analysis kernels (chrana.cu): frame observables with FP64 sums, d(s)/P(s)
// and the packed contact map in ccp's FP32 arithmetic.
// Reference behaviour followed: chrsim.cu:13-416, sugrid.cu:15-133,
// chrana.cu:13-84,372-512.

#include "kernels.hpp"

namespace ccpmi {
namespace {

__device__ __forceinline__ uint32_t wbits(float4 v) { return __float_as_uint(v.w); }

__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// ccp's Wang-Frenkel factor e (18 d^4 - 96 d^2 + 96) / d^8 (calc_ppf<DST>,
// chrsim.cu:173-185; calc_wpf, :79-88, with e = 1) in ccp's arithmetic: d is
// the FP32 length, d^2 its FP32 square, the polynomial and the division in
// FP64 (ccp's double literals promote them), d^8 the FP32 product
// ((d^2 d^2) d^2) d^2, and the factor rounded to FP32 before it scales the
// vector (vect.cuh's operator*(float, vec3f)).
__device__ __forceinline__ float ccp_wf(float e, float d) {
#pragma clang fp contract(off)
  const float d2 = d * d;
  const double x = (double)d2;
  const double poly = 18.0 * x * x - 96.0 * x + 96.0;
  const float d8 = d2 * d2 * d2 * d2;
  return (float)((double)e * poly / (double)d8);
}

// ccp's binding factor e_l 8/(3 lco^2) ((d/lco)^6 - 1) (calc_lpf<DST>,
// chrsim.cu:201-212): nd = d/lco and its square in FP32, the cube of the
// square in FP32, the rest in FP64 (lfac = e_l 8/(3 lco^2) in FP64).
__device__ __forceinline__ float ccp_lbs(double lfac, float d, float lco) {
#pragma clang fp contract(off)
  const float nd = d / lco;
  const float d2 = nd * nd;
  return (float)(lfac * ((double)(d2 * d2 * d2) - 1.0));
}
// Squared distance in one fixed rounding order, fma(dz, dz, fma(dy, dy,
// dx * dx)). Written as a sum of three products it leaves the compiler free
// to fuse either product of the first sum, and two inlined copies of the
// same pair term may be fused differently; a sum over a list would then
// depend on which copy took each entry.
__device__ __forceinline__ float dist2(float dx, float dy, float dz) { return fmaf(dz, dz, fmaf(dy, dy, dx * dx)); }

// the element at byte offset `off` of `base` (a 32-bit offset: the load takes
// the scalar base and the offset register as they are)
template <class T> __device__ __forceinline__ T at_byte(const T *base, uint32_t off) {
  return *(const T *)((const char *)base + off);
}

// ---------------------------------------------------------------- Philox4x32-10
// Salmon, Moraes, Dror, Shaw (SC'11): counter-based, one call per bead and
// step; no state array (ccp keeps a curandStatePhilox4_32_10 per bead).
__device__ __forceinline__ void philox10(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3, uint32_t k0,
                                         uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// ((x >> 8) + 1) * 2^-24: uniform in (0, 1], exact in FP32
__device__ __forceinline__ float u24(uint32_t x) { return (float)((x >> 8) + 1u) * 5.9604644775390625e-8f; }

// Normal triple (Box–Muller) with ccp's redraw of all three while any
// |z| > 5 (begin_iter, chrsim.cu:348-356). Counter {bead, step lo, step hi, attempt}.
__device__ __forceinline__ uint32_t noise3(uint32_t k0, uint32_t k1, uint32_t bead, uint64_t step, float &z0,
                                           float &z1, float &z2) {
  uint32_t a = 0;
  for (;; ++a) {
    uint32_t c0 = bead, c1 = (uint32_t)step, c2 = (uint32_t)(step >> 32), c3 = a;
    philox10(c0, c1, c2, c3, k0, k1);
    const float r01 = sqrtf(-2.0f * logf(u24(c0)));
    float s1, co1;
    sincospif(2.0f * u24(c1), &s1, &co1);
    const float r23 = sqrtf(-2.0f * logf(u24(c2)));
    const float co3 = cospif(2.0f * u24(c3));
    z0 = r01 * co1;
    z1 = r01 * s1;
    z2 = r23 * co3;
    if (!(fabsf(z0) > 5.0f || fabsf(z1) > 5.0f || fabsf(z2) > 5.0f) || a == 0xFFFFu) break;
  }
  return a;
}

// ---------------------------------------------------------------- wave helpers
// v of another lane through a DPP move: quad_perm [1,0,3,2] (lane ^ 1),
// [2,3,0,1] (lane ^ 2), the half-row mirror (lane 7 - i within eight), row
// shifts right by n (lane i - n within its 16-lane row) and the row
// broadcasts (lane 15 of each row into the next row, lane 31 into rows 2
// and 3). Lanes outside ROWS (16-lane rows) or BANKS (4-lane groups of each
// row), or with no source lane, get 0.
constexpr int DPP_XOR1 = 0xB1, DPP_XOR2 = 0x4E, DPP_MIRROR8 = 0x141, DPP_SHR1 = 0x111, DPP_SHR2 = 0x112,
              DPP_SHR4 = 0x114, DPP_SHR8 = 0x118, DPP_BCAST15 = 0x142, DPP_BCAST31 = 0x143;
template <int CTRL, int BANKS = 0xF, int ROWS = 0xF> __device__ __forceinline__ uint32_t dpp_u(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, BANKS, true);
}
template <int CTRL> __device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float((int)dpp_u<CTRL>(__float_as_uint(v)));
}

// Inclusive prefix over the 64 lanes (all active): within each 16-lane row
// by shifts of 1, 2, 4, 8, then the rows' totals carried by the broadcasts.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  v += dpp_u<DPP_SHR1>(v);
  v += dpp_u<DPP_SHR2>(v);
  v += dpp_u<DPP_SHR4>(v);
  v += dpp_u<DPP_SHR8>(v);
  v += dpp_u<DPP_BCAST15, 0xF, 0xA>(v);
  v += dpp_u<DPP_BCAST31, 0xF, 0xC>(v);
  return v;
}

// sum over the 64 lanes (all active), the same in every lane
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan(v), 63);
}

// maximum over the 64 lanes (all active), the same in every lane: the
// steps of wave_incl_scan with max for +
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
  v = max(v, dpp_u<DPP_SHR1>(v));
  v = max(v, dpp_u<DPP_SHR2>(v));
  v = max(v, dpp_u<DPP_SHR4>(v));
  v = max(v, dpp_u<DPP_SHR8>(v));
  v = max(v, dpp_u<DPP_BCAST15, 0xF, 0xA>(v));
  v = max(v, dpp_u<DPP_BCAST31, 0xF, 0xC>(v));
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// atomicAdd(base[key], 1) for every active lane, one atomic per distinct key
// in the wavefront (ballot match loop)
__device__ __forceinline__ void agg_add_noret(uint32_t *base, uint32_t key) {
  const int lane = __lane_id();
  unsigned long long todo = __ballot(1);
  while (todo) {
    const int leader = __ffsll(todo) - 1;
    const uint32_t lkey = __shfl(key, leader);
    const unsigned long long same = __ballot(key == lkey) & todo;
    if (lane == leader) atomicAdd(base + lkey, (uint32_t)__popcll(same));
    todo &= ~same;
  }
}

// ---------------------------------------------------------------- grid
// Cell coordinate along one axis: floor(x / csl) + cps/2 as ccp's
// calc_indexes (sugrid.cu:26-34), clamped into the grid (ccp has no clamp).
// Step grids multiply by 1/csl instead of dividing: every lookup of a grid
// uses the same rule, so binning and neighbour scans stay consistent.
__device__ __forceinline__ int cell_coord(float x, const GridGeom &g, bool &outside) {
  int c = (int)floorf(g.icsl != 0.0f ? x * g.icsl : x / g.csl) + g.half;
  if (c < 0) { c = 0; outside = true; }
  if (c >= g.cps) { c = g.cps - 1; outside = true; }
  return c;
}

// The bucket entry is the member's position with its index in the w bits.
// hblocksum (optional) counts the het members per scan block, for the order
// pass's type-ordered work list.
__device__ __forceinline__ void bin_into(const GridGeom &g, uint32_t *count, uint32_t *blocksum, float4 *bucket,
                                         uint32_t *flags, float4 v, uint32_t *hblocksum = nullptr) {
  bool outside = false;
  const int gx = cell_coord(v.x, g, outside), gy = cell_coord(v.y, g, outside), gz = cell_coord(v.z, g, outside);
  const uint32_t key = (uint32_t)(g.coff + (gz * g.cps + gy) * g.cps + gx);
  // beads of one wavefront mostly sit in distinct cells: plain atomics
  const uint32_t slot = atomicAdd(count + key, 1u);
  // the block totals and the flag need no slot: done while it returns
  agg_add_noret(blocksum, key / SCAN_BLOCK);
  if (hblocksum && (wbits(v) & W_HET)) agg_add_noret(hblocksum, key / SCAN_BLOCK);
  if (outside) atomicAdd(flags + FLAG_OUTSIDE, 1u);
  if (slot < (uint32_t)g.cap) bucket[(size_t)key * g.cap + slot] = v;
}

// Bin a new bead position into the next fine grid and, for het beads, the
// next het grid (the het-grid atomics run on the het lanes only).
__device__ __forceinline__ void bin_bead(const KParams &p, const KPtrs &q, uint32_t w, float x, float y, float z) {
  const float4 v = make_float4(x, y, z, __uint_as_float(w)); // w carries the bead index
  bin_into(p.gf, q.fcount, q.fblocksum, q.fbucket, q.flags, v, q.fhblocksum);
  if (p.gh.n_c > 0 && (w & W_HET)) bin_into(p.gh, q.hcount, q.hblocksum, q.hbucket, q.flags, v);
  if (!(isfinite(x) && isfinite(y) && isfinite(z))) atomicAdd(q.flags + FLAG_NONFINITE, 1u);
}

// Slot ranges of the 9 neighbour rows (x-1..x+1 of rows y+dy, z+dz) of the
// cell holding (x, y, z); rows outside the grid get an empty range. All 18
// start lookups are issued together.
struct Rows {
  uint32_t lo[9], hi[9];
  int c0[9], c1[9];
};

__device__ __forceinline__ void neighbour_rows(const GridGeom &g, const uint32_t *__restrict__ cstart, float x,
                                               float y, float z, Rows &R) {
  bool outside = false;
  const int cx = cell_coord(x, g, outside), cy = cell_coord(y, g, outside), cz = cell_coord(z, g, outside);
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    const int zz = cz + r / 3 - 1, yy = cy + r % 3 - 1;
    const bool ok = zz >= 0 && zz < g.cps && yy >= 0 && yy < g.cps;
    const int row = g.coff + (zz * g.cps + yy) * g.cps;
    R.c0[r] = ok ? row + max(cx - 1, 0) : 0;
    R.c1[r] = ok ? row + min(cx + 1, g.cps - 1) + 1 : 0;
    R.lo[r] = cstart[R.c0[r]];
    R.hi[r] = cstart[R.c1[r]];
  }
}

// ---------------------------------------------------------------- forces
// Short-range pair force on bead i from entry o (every pair type), with
// |i - j| <= 1 skipped (calc_cell_srf, chrsim.cu:250).
// CCP: Wang-Frenkel e_p (18 d^4 - 96 d^2 + 96)/d^8 (calc_ppf<DST>,
// chrsim.cu:173-185) for d <= rco, the cutoff every pair type has.
// ICG: soft repulsion 128 (3 rco - 3 d) for d <= rco (calc_ppf<ICG>, :188-198).
// LJ: WCA for d < 2^(1/6) sigma.
template <int MODE>
__device__ __forceinline__ void pair_short(const KParams &p, float rx, float ry, float rz, uint32_t wi, float4 o,
                                           float &fx, float &fy, float &fz) {
  const uint32_t wj = wbits(o);
  const uint32_t i = wi & ID_MASK, j = wj & ID_MASK;
  const float dx = rx - o.x, dy = ry - o.y, dz = rz - o.z;
  const float d2 = dist2(dx, dy, dz);
  const bool nb = (j > i + 1u) || (i > j + 1u);
  float fac;
  if constexpr (MODE == MODE_CCP) {
    const float d = sqrtf(d2);
    fac = (nb && d <= p.rco) ? ccp_wf(p.e_p, d) : 0.0f;
  } else if constexpr (MODE == MODE_ICG) {
    const float d = sqrtf(d2);
    fac = 128.0f * (3.0f * p.rco - 3.0f * d);
    fac = (nb && d <= p.rco) ? fac : 0.0f;
  } else {
    const float u = fast_rcp(d2);
    const float s2 = u;
    const float s6 = s2 * s2 * s2;
    fac = p.eps_wca * 24.0f * s6 * (2.0f * s6 - 1.0f) * u;
    fac = (nb && d2 < p.rm2) ? fac : 0.0f;
  }
  fx += fac * dx;
  fy += fac * dy;
  fz += fac * dz;
}

// Long-range het-het pair force (both beads het by construction).
// CCP: the Wang-Frenkel force of A-A pairs for rco < d <= aco.
// LJ: the LJ tail eps_het for 2^(1/6) sigma <= d < rc_lj.
template <int MODE>
__device__ __forceinline__ void pair_tail(const KParams &p, float rx, float ry, float rz, uint32_t wi, float4 o,
                                          float &fx, float &fy, float &fz) {
  const uint32_t wj = wbits(o);
  const uint32_t i = wi & ID_MASK, j = wj & ID_MASK;
  const float dx = rx - o.x, dy = ry - o.y, dz = rz - o.z;
  const float d2 = dist2(dx, dy, dz);
  const bool nb = (j > i + 1u) || (i > j + 1u);
  float fac;
  if constexpr (MODE == MODE_CCP) {
    const float d = sqrtf(d2);
    fac = (nb && d > p.rco && d <= p.aco) ? ccp_wf(p.e_p, d) : 0.0f;
  } else {
    const float u = fast_rcp(d2);
    const float s2 = u;
    const float s6 = s2 * s2 * s2;
    fac = p.eps_het * 24.0f * s6 * (2.0f * s6 - 1.0f) * u;
    fac = (nb && d2 >= p.rm2 && d2 < p.rc_lj2) ? fac : 0.0f;
  }
  fx += fac * dx;
  fy += fac * dy;
  fz += fac * dz;
}

// LBS binding e_l 8/(3 lco^2) ((d/lco)^6 - 1) (calc_lpf<DST>, chrsim.cu:201-212).
__device__ __forceinline__ void lbs_force(const KParams &p, float rx, float ry, float rz, float4 o, float &fx,
                                          float &fy, float &fz) {
  const float dx = rx - o.x, dy = ry - o.y, dz = rz - o.z;
  const float d = sqrtf(dist2(dx, dy, dz));
  const float fac = d <= p.lco ? ccp_lbs(p.lfac_d, d, p.lco) : 0.0f;
  fx += fac * dx;
  fy += fac * dy;
  fz += fac * dz;
}

// Bonds and bending on bead i (calc_bf, chrsim.cu:13-76); bond existence
// comes from the W_B* bits, positions from the bead-ordered array.
// nb(k): the position of bead i + k (k = -2, -1, 1, 2), read only where the
// chain has that neighbour
template <class NB>
__device__ __forceinline__ void bonded_terms(const KParams &p, NB nb, uint32_t w, float rx, float ry, float rz,
                                             float &fx, float &fy, float &fz) {
  float v0x = 0, v0y = 0, v0z = 0, v1x = 0, v1y = 0, v1z = 0, v2x = 0, v2y = 0, v2z = 0, v3x = 0, v3y = 0, v3z = 0;
  float il0 = 0, il1 = 0, il2 = 0, il3 = 0;
  if (w & W_B1) {
    const float4 a = nb(-1);
    v1x = rx - a.x; v1y = ry - a.y; v1z = rz - a.z;
    il1 = rsqrtf(v1x * v1x + v1y * v1y + v1z * v1z);
    if (w & W_B0) {
      const float4 b = nb(-2);
      v0x = a.x - b.x; v0y = a.y - b.y; v0z = a.z - b.z;
      il0 = rsqrtf(v0x * v0x + v0y * v0y + v0z * v0z);
    }
  }
  if (w & W_B2) {
    const float4 a = nb(1);
    v2x = a.x - rx; v2y = a.y - ry; v2z = a.z - rz;
    il2 = rsqrtf(v2x * v2x + v2y * v2y + v2z * v2z);
    if (w & W_B3) {
      const float4 b = nb(2);
      v3x = b.x - a.x; v3y = b.y - a.y; v3z = b.z - a.z;
      il3 = rsqrtf(v3x * v3x + v3y * v3y + v3z * v3z);
    }
  }
  if (p.bond_type == 0) {
    const float c2 = p.k_e * (1.0f - p.l_0 * il2), c1 = p.k_e * (1.0f - p.l_0 * il1);
    fx += c2 * v2x - c1 * v1x;
    fy += c2 * v2y - c1 * v1y;
    fz += c2 * v2z - c1 * v1z;
  } else {
    // FENE K b / (1 - b^2/R0^2) plus WCA between bonded beads
    if (w & W_B2) {
      const float d2 = v2x * v2x + v2y * v2y + v2z * v2z;
      float c = p.fene_k / fmaxf(1.0f - d2 / p.fene_r02, 0.01f);
      if (d2 < p.rm2) {
        const float u = fast_rcp(d2), s2 = u, s6 = s2 * s2 * s2;
        c -= p.eps_wca * 24.0f * s6 * (2.0f * s6 - 1.0f) * u;
      }
      fx += c * v2x; fy += c * v2y; fz += c * v2z;
    }
    if (w & W_B1) {
      const float d2 = v1x * v1x + v1y * v1y + v1z * v1z;
      float c = -p.fene_k / fmaxf(1.0f - d2 / p.fene_r02, 0.01f);
      if (d2 < p.rm2) {
        const float u = fast_rcp(d2), s2 = u, s6 = s2 * s2 * s2;
        c += p.eps_wca * 24.0f * s6 * (2.0f * s6 - 1.0f) * u;
      }
      fx += c * v1x; fy += c * v1y; fz += c * v1z;
    }
  }
  if (p.bend_all || (w & W_HET)) {
    const float cs0 = (v1x * v0x + v1y * v0y + v1z * v0z) * il1 * il0;
    const float cs1 = (v2x * v1x + v2y * v1y + v2z * v1z) * il2 * il1;
    const float cs2 = (v3x * v2x + v3y * v2y + v3z * v2z) * il3 * il2;
    const float a0 = il1 * il0, a1 = il1 * il2, a3 = il2 * il3;
    const float q1 = il1 * il1, q2 = il2 * il2;
    // k_b [ a0 v0 - cs0 q1 v1 + a1 v2 - cs1 q1 v1 - a1 v1 + cs1 q2 v2 - a3 v3 + cs2 q2 v2 ]
    const float c0 = p.k_b * a0;
    const float c1 = -p.k_b * (cs0 * q1 + cs1 * q1 + a1);
    const float c2 = p.k_b * (a1 + cs1 * q2 + cs2 * q2);
    const float c3 = -p.k_b * a3;
    fx += c0 * v0x + c1 * v1x + c2 * v2x + c3 * v3x;
    fy += c0 * v0y + c1 * v1y + c2 * v2y + c3 * v3y;
    fz += c0 * v0z + c1 * v1z + c2 * v2z + c3 * v3z;
  }
}

__device__ __forceinline__ void bonded_force(const KParams &p, const float4 *__restrict__ pos, uint32_t i,
                                             uint32_t w, float rx, float ry, float rz, float &fx, float &fy,
                                             float &fz) {
  bonded_terms(p, [&](int k) { return pos[i + k]; }, w, rx, ry, rz, fx, fy, fz);
}

// Confinement (calc_cf<DST>, chrsim.cu:91-144, sphere + opening + bleb;
// calc_cf<ICG>, :147-170, sphere only) with the Wang-Frenkel wall
// (calc_wpf, :79-88) or, for LJ, WCA plus a lamina tail on LAD beads.
template <int MODE>
__device__ __forceinline__ void wall_force(const KParams &p, float rx, float ry, float rz, uint32_t w, float &fx,
                                           float &fy, float &fz) {
  float vx, vy, vz;
  if constexpr (MODE == MODE_ICG) {
    const float d_r = sqrtf(rx * rx + ry * ry + rz * rz);
    const float s = -(p.R_n / d_r - 1.0f);
    vx = rx * s; vy = ry * s; vz = rz * s;
  } else {
    if (rz < p.nod) {
      float d_r = sqrtf(rx * rx + ry * ry + rz * rz);
      if (rz / d_r < p.noc) {
        const float s = -(p.R_n / d_r - 1.0f);
        vx = rx * s; vy = ry * s; vz = rz * s;
      } else {
        d_r = sqrtf(rx * rx + ry * ry);
        const float s = -(p.R_o / d_r - 1.0f);
        vx = rx * s; vy = ry * s; vz = rz - p.nod;
      }
    } else {
      const float qz = rz - (p.nod + p.bod);
      float d_r = sqrtf(rx * rx + ry * ry + qz * qz);
      if (-qz / d_r < p.boc) {
        const float s = -(p.R_b / d_r - 1.0f);
        vx = rx * s; vy = ry * s; vz = qz * s;
      } else {
        d_r = sqrtf(rx * rx + ry * ry);
        const float s = -(p.R_o / d_r - 1.0f);
        vx = rx * s; vy = ry * s; vz = qz + p.bod;
      }
    }
  }
  const float dw2 = vx * vx + vy * vy + vz * vz;
  if (!isfinite(dw2)) return;
  float fac;
  if constexpr (MODE != MODE_LJ) {
    const float d = sqrtf(dw2);
    if (d > p.rco) return;
    fac = ccp_wf(1.0f, d);
  } else {
    const float u = fast_rcp(dw2);
    const float s2 = u, s6 = s2 * s2 * s2;
    const float base = 24.0f * s6 * (2.0f * s6 - 1.0f) * u;
    if (dw2 < p.rm2) fac = p.eps_wall * base;
    else if ((w & W_LAD) && dw2 < p.rc_lam2) fac = p.eps_lam * base;
    else return;
  }
  fx += fac * vx;
  fy += fac * vy;
  fz += fac * vz;
}

// ---------------------------------------------------------------- kernels
// Complete force of a non-bonded pair (|i - j| > 1): ccp's calc_ppf<DST>
// (Wang-Frenkel, cut at aco for het-het pairs and at rco otherwise,
// chrsim.cu:173-185), calc_ppf<ICG> (soft repulsion within rco, :188-198),
// or WCA plus the het-het LJ tail.
template <int MODE>
__device__ __forceinline__ void pair_full(const KParams &p, float rx, float ry, float rz, uint32_t wi, float4 o,
                                          float &fx, float &fy, float &fz) {
  const uint32_t wj = wbits(o);
  const uint32_t i = wi & ID_MASK, j = wj & ID_MASK;
  const float dx = rx - o.x, dy = ry - o.y, dz = rz - o.z;
  const float d2 = dist2(dx, dy, dz);
  // |i - j| > 1 as one unsigned compare: j - (i - 1) wraps past 2 unless
  // j is i - 1, i or i + 1
  const bool nb = j - (i - 1u) > 2u;
  const bool hhi = (wi & wj & W_HET) != 0u;
  float fac;
  if constexpr (MODE == MODE_CCP) {
    const float d = sqrtf(d2);
    fac = (nb && d <= (hhi ? p.aco : p.rco)) ? ccp_wf(p.e_p, d) : 0.0f;
  } else if constexpr (MODE == MODE_ICG) {
    const float d = sqrtf(d2);
    fac = 128.0f * (3.0f * p.rco - 3.0f * d);
    fac = (nb && d <= p.rco) ? fac : 0.0f;
  } else {
    // sigma = 1: 24 eps (2 r^-12 - r^-6) / r^2, WCA within 2^(1/6) for
    // every pair, the LJ tail beyond it for het-het pairs. The conditions
    // are combined with & and | (no short circuit), so they stay lane masks
    // instead of 0/1 values.
    const float u = fast_rcp(d2);
    const float s6 = u * u * u;
    const float base = s6 * fmaf(48.0f, s6, -24.0f) * u;
    const bool wca = d2 < p.rm2;
    const bool on = nb & (wca | (hhi & (d2 < p.rc_lj2)));
    fac = on ? (wca ? p.eps_wca : p.eps_het) * base : 0.0f;
  }
  fx += fac * dx;
  fy += fac * dy;
  fz += fac * dz;
}

// Per-bead operands of pair_full held in registers across a bead's candidate
// loop. i - 1 is hidden from the compiler, which would otherwise rewrite
// j - (i - 1) as (j - i) + 1, two operations per pair; both LJ well depths sit
// in vector registers, since a select takes at most one scalar operand.
struct PairCtx {
  float rx, ry, rz, e_wca, e_het;
  uint32_t im1, het;
  uint32_t nim1; // 1 - i
};
__device__ __forceinline__ PairCtx pair_ctx(const KParams &p, float4 me) {
  PairCtx c;
  const uint32_t w = wbits(me);
  c.rx = me.x;
  c.ry = me.y;
  c.rz = me.z;
  c.im1 = (w & ID_MASK) - 1u;
  c.nim1 = 0u - c.im1;
  c.het = w & W_HET;
  c.e_wca = p.eps_wca;
  c.e_het = p.eps_het;
  asm("" : "+v"(c.im1), "+v"(c.nim1), "+v"(c.e_wca), "+v"(c.e_het));
  return c;
}

// (the partner's index) - (i - 1), the index taken from the low 24 bits of
// the partner's word by the multiply-add itself
__device__ __forceinline__ uint32_t index_gap(uint32_t wj, uint32_t nim1) {
  static_assert(ID_BITS == 24, "v_mad_u32_u24 reads a 24-bit index");
  uint32_t r;
  asm("v_mad_u32_u24 %0, %1, 1, %2" : "=v"(r) : "v"(wj), "v"(nim1));
  return r;
}

// Kind of bead of a pair term, when the caller knows it: any, a bead with
// no het-het pairs (non-het), or a het bead (every het partner is a
// het-het pair).
enum BeadKind : int { BEAD_ANY = 0, BEAD_SHORT = 1, BEAD_HET = 2 };

// pair_full with the bead's operands from a PairCtx: the same operations in
// the same order, so the same bits. For a known kind the het-het test is
// one sign test (het bead) or none (non-het bead).
template <int MODE, int KIND = BEAD_ANY>
__device__ __forceinline__ void pair_term(const KParams &p, const PairCtx &c, float4 o, float &fx, float &fy,
                                          float &fz) {
  static_assert(W_HET == 0x80000000u, "the het flag is the sign bit");
  const uint32_t wj = wbits(o);
  const float dx = c.rx - o.x, dy = c.ry - o.y, dz = c.rz - o.z;
  const float d2 = dist2(dx, dy, dz);
  const bool nb = index_gap(wj, c.nim1) > 2u;
  const bool hhi = KIND == BEAD_SHORT ? false : KIND == BEAD_HET ? (int32_t)wj < 0 : (wj & c.het) != 0u;
  float fac;
  if constexpr (MODE == MODE_CCP) {
    const float d = sqrtf(d2);
    fac = (nb & (d <= (hhi ? p.aco : p.rco))) ? ccp_wf(p.e_p, d) : 0.0f;
  } else if constexpr (MODE == MODE_ICG) {
    const float d = sqrtf(d2);
    fac = 128.0f * (3.0f * p.rco - 3.0f * d);
    fac = (nb & (d <= p.rco)) ? fac : 0.0f;
  } else {
    // the force is added only where the pair is on: a wavefront in which no
    // lane's candidate is within its cutoff skips the whole block
    const bool wca = d2 < p.rm2;
    const bool on = nb & (wca | (hhi & (d2 < p.rc_lj2)));
    if (on) {
      const float u = fast_rcp(d2);
      const float s6 = u * u * u;
      const float base = s6 * fmaf(48.0f, s6, -24.0f) * u;
      const float f = (KIND == BEAD_SHORT ? c.e_wca : wca ? c.e_wca : c.e_het) * base;
      fx += f * dx;
      fy += f * dy;
      fz += f * dz;
    }
    return;
  }
  fx += fac * dx;
  fy += fac * dy;
  fz += fac * dz;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2. Renumber them so XCD x runs one contiguous run of tiles: its beads,
// and the neighbour rows they read, then stay in that XCD's L2.
constexpr int N_XCD = 8;
// block b of a run of nb blocks whose first block sits on XCD 0
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int x = b % N_XCD, k = b / N_XCD, q = nb / N_XCD, rem = nb % N_XCD;
  return (x < rem ? x * (q + 1) : rem * (q + 1) + (x - rem) * q) + k;
}
__device__ __forceinline__ int xcd_block() { return xcd_remap((int)blockIdx.x, (int)gridDim.x); }

// The 2R+1 neighbour rows (y-R..y+R, x-R..x+R) of plane z + dz around the
// cell holding (x, y, z); rows outside the grid get an empty range. Reach R
// = 1 covers ccp's 27 cells; R = 2 covers 125 cells of half the side (the
// same cutoff sphere with fewer candidates).
template <int R> struct RowsR {
  uint32_t lo[2 * R + 1], hi[2 * R + 1];
  int c0[2 * R + 1], c1[2 * R + 1];
};

template <int R>
__device__ __forceinline__ void plane_rows(const GridGeom &g, const uint32_t *__restrict__ cstart, float x, float y,
                                           float z, int dz, RowsR<R> &Q) {
  bool outside = false;
  const int cx = cell_coord(x, g, outside), cy = cell_coord(y, g, outside), cz = cell_coord(z, g, outside);
  const int zz = cz + dz;
#pragma unroll
  for (int r = 0; r < 2 * R + 1; ++r) {
    const int yy = cy + r - R;
    const bool ok = zz >= 0 && zz < g.cps && yy >= 0 && yy < g.cps;
    const int row = g.coff + (zz * g.cps + yy) * g.cps;
    Q.c0[r] = ok ? row + max(cx - R, 0) : 0;
    Q.c1[r] = ok ? row + min(cx + R, g.cps - 1) + 1 : 0;
    Q.lo[r] = cstart[Q.c0[r]];
    Q.hi[r] = cstart[Q.c1[r]];
  }
}

// Candidates of the rows, four at a time (independent loads in flight).
// PASS 0: every pair on the all-bead grid (full cutoffs, or the short
// cutoff when split); PASS 1: het-het tail on the het grid.
template <int MODE, int PASS, bool SPLIT, int R>
__device__ __forceinline__ void scan_pass(const KParams &p, const float4 *__restrict__ spos, const RowsR<R> &Q,
                                          float rx, float ry, float rz, uint32_t w, float &fx, float &fy, float &fz) {
#pragma unroll
  for (int r = 0; r < 2 * R + 1; ++r) {
    uint32_t k = Q.lo[r];
    const uint32_t h = Q.hi[r];
    for (; k + 4u <= h; k += 4u) {
      const float4 a = spos[k], b = spos[k + 1], c = spos[k + 2], d = spos[k + 3];
      if constexpr (PASS == 1) {
        pair_tail<MODE>(p, rx, ry, rz, w, a, fx, fy, fz);
        pair_tail<MODE>(p, rx, ry, rz, w, b, fx, fy, fz);
        pair_tail<MODE>(p, rx, ry, rz, w, c, fx, fy, fz);
        pair_tail<MODE>(p, rx, ry, rz, w, d, fx, fy, fz);
      } else if constexpr (SPLIT) {
        pair_short<MODE>(p, rx, ry, rz, w, a, fx, fy, fz);
        pair_short<MODE>(p, rx, ry, rz, w, b, fx, fy, fz);
        pair_short<MODE>(p, rx, ry, rz, w, c, fx, fy, fz);
        pair_short<MODE>(p, rx, ry, rz, w, d, fx, fy, fz);
      } else {
        pair_full<MODE>(p, rx, ry, rz, w, a, fx, fy, fz);
        pair_full<MODE>(p, rx, ry, rz, w, b, fx, fy, fz);
        pair_full<MODE>(p, rx, ry, rz, w, c, fx, fy, fz);
        pair_full<MODE>(p, rx, ry, rz, w, d, fx, fy, fz);
      }
    }
    for (; k < h; ++k) {
      if constexpr (PASS == 1) pair_tail<MODE>(p, rx, ry, rz, w, spos[k], fx, fy, fz);
      else if constexpr (SPLIT) pair_short<MODE>(p, rx, ry, rz, w, spos[k], fx, fy, fz);
      else pair_full<MODE>(p, rx, ry, rz, w, spos[k], fx, fy, fz);
    }
  }
}

// LANES = 2R+1: one lane per z-plane of the neighbourhood, the partial
// forces summed with shuffles (R = 1: 21 beads per wavefront, lane 63
// idles; R = 2: 12 beads, lanes 60-63 idle), so N beads give LANES N / 63
// wavefronts to hide the gather latency. LANES = 1: one lane scans every
// plane.
template <int LANES> struct Lanes {
  static constexpr int BEADS_PER_WAVE = 64 / LANES;
  static constexpr int BEADS_PER_BLOCK = BEADS_PER_WAVE * (BEAD_BLOCK / 64);
};

// System of cell-ordered slot t in a batch: the last s with sys_slot[s] <= t.
__device__ __forceinline__ int system_of(const uint32_t *__restrict__ sys_slot, int n_sys, uint32_t t) {
  int lo = 0, hi = n_sys - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sys_slot[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Constants of the batch system holding slot t.
__device__ __forceinline__ void load_system(KParams &p, const KPtrs &q, uint32_t t) {
  const SysParams &S = q.sys[system_of(q.sys_slot, q.n_sys, t)];
  p.R_n = S.R_n; p.R_o = S.R_o; p.R_b = S.R_b;
  p.noc = S.noc; p.boc = S.boc; p.nod = S.nod; p.bod = S.bod;
  p.eps_het = S.eps_het; p.eps_lam = S.eps_lam;
  p.gf = S.g;
  p.frozen = S.frozen;
}

// Unit normal triple of bead g (index over all systems) for the job's step.
__device__ __forceinline__ void fill_noise(const NoiseJob &nj, uint32_t g) {
  uint32_t k0 = nj.seed_lo, k1 = nj.seed_hi, local = g;
  if (nj.sys) {
    const SysParams &S = nj.sys[system_of(nj.sys_slot, nj.n_sys, g)];
    k0 = S.seed_lo;
    k1 = S.seed_hi;
    local = g - S.slot;
  }
  float z0, z1, z2;
  const uint32_t a = noise3(k0, k1, local, *nj.step + nj.ahead, z0, z1, z2);
  nj.z[g] = make_float4(z0, z1, z2, __uint_as_float(a));
}

// Bonded + wall force of bead g for the next force pass (one lane per bead).
// The bead and its four chain neighbours (indices clamped to [0, n), read
// whether or not the chain has them), loaded together so that they share one
// wait; bonded_terms uses only the ones it needs.
struct BondNbrs {
  float4 me, m2, m1, p1, p2;
};
__device__ __forceinline__ BondNbrs load_bond_nbrs(const BondJob &bj, uint32_t g) {
  BondNbrs b;
  b.me = bj.pos[g];
  b.m2 = bj.pos[g >= 2u ? g - 2u : g];
  b.m1 = bj.pos[g >= 1u ? g - 1u : g];
  b.p1 = bj.pos[g + 1u < bj.n ? g + 1u : g];
  b.p2 = bj.pos[g + 2u < bj.n ? g + 2u : g];
  return b;
}
__device__ __forceinline__ void fill_bond(KParams p, const BondJob &bj, uint32_t g, const BondNbrs &b) {
  if (bj.sys) {
    const SysParams &S = bj.sys[system_of(bj.sys_slot, bj.n_sys, g)];
    p.R_n = S.R_n; p.R_o = S.R_o; p.R_b = S.R_b;
    p.noc = S.noc; p.boc = S.boc; p.nod = S.nod; p.bod = S.bod;
    p.eps_het = S.eps_het; p.eps_lam = S.eps_lam;
  }
  const float4 me = b.me;
  const uint32_t w = wbits(me);
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  bonded_terms(p, [&](int k) { return k == -2 ? b.m2 : k == -1 ? b.m1 : k == 1 ? b.p1 : b.p2; }, w, me.x, me.y,
               me.z, fx, fy, fz);
  if (bj.mode == MODE_LJ) wall_force<MODE_LJ>(p, me.x, me.y, me.z, w, fx, fy, fz);
  else if (bj.mode == MODE_ICG) wall_force<MODE_ICG>(p, me.x, me.y, me.z, w, fx, fy, fz);
  else wall_force<MODE_CCP>(p, me.x, me.y, me.z, w, fx, fy, fz);
  bj.out[g] = make_float4(fx, fy, fz, 0.0f);
}

// Force evaluation of one stage, beads in grid order.
// STAGE 0: write F(pos) by bead index. STAGE 1: Heun predictor
// (exec_RK_1, chrsim.cu:364-388) er = r + f dt + rn, keeps
// r + f dt/2 + rn. STAGE 2: corrector (exec_RK_2, :391-416)
// r = (r + f dt/2 + rn) + ef dt/2. Stages 1 and 2 bin their output.
// BATCH: the slot's system supplies the nucleus geometry, eps_het, eps_lam,
// the noise key and the grid (its cells are a range of the shared ones);
// the noise counter uses the bead's index within its system, so a system
// evolves exactly as it would alone with the same key.
template <int MODE, int STAGE, bool SPLIT, bool BATCH, int R, int LANES>
__global__ __launch_bounds__(BEAD_BLOCK) void k_force(KParams p, KPtrs q) {
  static_assert(LANES == 1 || LANES == 2 * R + 1, "one lane per bead or one per plane");
  constexpr int BPW = Lanes<LANES>::BEADS_PER_WAVE;
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  const int part = LANES > 1 ? lane % LANES : 0, slot = LANES > 1 ? lane / LANES : lane;
  const int t = (xcd_block() * (BEAD_BLOCK / 64) + wave) * BPW + slot;
  const bool in_range = slot < BPW && t < p.N;
  const float4 me = in_range ? q.fspos[t] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const uint32_t w = wbits(me);
  const uint32_t i = w & ID_MASK;
  const bool active = in_range && i < (uint32_t)p.N; // a sentinel slot holds no bead
  const float rx = me.x, ry = me.y, rz = me.z;
  if constexpr (BATCH) load_system(p, q, active ? (uint32_t)t : 0u);

  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  if (active && part == 0) {
    bonded_force(p, q.pos, i, w, rx, ry, rz, fx, fy, fz);
    wall_force<MODE>(p, rx, ry, rz, w, fx, fy, fz);
  }
#pragma unroll
  for (int pl = 0; pl < (LANES > 1 ? 1 : 2 * R + 1); ++pl) {
    const int dz = (LANES > 1 ? part : pl) - R;
    RowsR<R> Q;
    plane_rows<R>(p.gf, q.fcstart, rx, ry, rz, dz, Q);
    if (!active)
#pragma unroll
      for (int r = 0; r < 2 * R + 1; ++r) Q.hi[r] = Q.lo[r];
    scan_pass<MODE, 0, SPLIT, R>(p, q.fspos, Q, rx, ry, rz, w, fx, fy, fz);
    if (!BATCH && MODE != MODE_ICG && active && p.n_l > 0 && (w & W_LAD)) {
#pragma unroll
      for (int r = 0; r < 2 * R + 1; ++r) {
        const uint32_t llo = q.lstart[Q.c0[r]], lhi = q.lstart[Q.c1[r]];
        for (uint32_t k = llo; k < lhi; ++k) lbs_force(p, rx, ry, rz, q.lspos[k], fx, fy, fz);
      }
    }
    if constexpr (SPLIT && MODE != MODE_ICG) {
      if (active && (w & W_HET)) {
        RowsR<1> H;
        plane_rows<1>(p.gh, q.hcstart, rx, ry, rz, dz, H);
        scan_pass<MODE, 1, true, 1>(p, q.hspos, H, rx, ry, rz, w, fx, fy, fz);
      }
    }
  }
  if constexpr (LANES > 1) {
    // sum the plane parts into lane part 0 (lane + k stays in the wave; the
    // idle top lanes contribute zero)
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
#pragma unroll
    for (int k = 1; k < LANES; ++k) {
      const int src = min(lane + k, 63);
      ax += __shfl(fx, src);
      ay += __shfl(fy, src);
      az += __shfl(fz, src);
    }
    if (!active || part != 0) return;
    fx += ax;
    fy += ay;
    fz += az;
  } else {
    if (!active) return;
  }
  if (p.frozen) fx = fy = fz = 0.0f; // with no noise (relaxation) the bead stays put
  if (STAGE != 0 && p.fcap2 > 0.0f) {
    // optional safeguard of the WCA/LJ set: |F| <= fcap
    const float m2 = fx * fx + fy * fy + fz * fz;
    if (m2 > p.fcap2) {
      const float sc = sqrtf(p.fcap2 / m2);
      fx *= sc; fy *= sc; fz *= sc;
    }
  }

  if constexpr (STAGE == 0) {
    q.out0[i] = make_float4(fx, fy, fz, 0.0f);
  } else if constexpr (STAGE == 1) {
    float nx = 0.0f, ny = 0.0f, nz = 0.0f;
    if (p.sd > 0.0f) {
      const float4 z = q.noise[i];
      nx = p.sd * z.x; ny = p.sd * z.y; nz = p.sd * z.z;
    }
    const float ex = rx + fx * p.dt + nx, ey = ry + fy * p.dt + ny, ez = rz + fz * p.dt + nz;
    q.out0[i] = make_float4(ex, ey, ez, me.w);
    q.pstar[i] = make_float4(rx + fx * p.half_dt + nx, ry + fy * p.half_dt + ny, rz + fz * p.half_dt + nz, 0.0f);
    bin_bead(p, q, w, ex, ey, ez);
  } else {
    const float4 ps = q.pstar[i];
    const float nx = ps.x + fx * p.half_dt, ny = ps.y + fy * p.half_dt, nz = ps.z + fz * p.half_dt;
    q.out0[i] = make_float4(nx, ny, nz, me.w);
    bin_bead(p, q, w, nx, ny, nz);
  }
}

// The per-bead input of the stage update: the step's unit normals (stage 1,
// zero without noise) or r + f dt/2 + noise (stage 2). Kernels may load it
// early so that its latency hides behind the pair sum.
template <int STAGE> __device__ __forceinline__ float4 stage_input(const KParams &p, const KPtrs &q, uint32_t i) {
  if constexpr (STAGE == 1) return p.sd > 0.0f ? q.noise[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  else if constexpr (STAGE == 2) return q.pstar[i];
  else return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Force cap and the stage update of bead i from its total force (the
// epilogue of the force kernels): STAGE 0 writes the force, STAGE 1 the Heun
// predictor and r + f dt/2 + noise, STAGE 2 the corrector; both bin the new
// position. `in` is stage_input<STAGE>.
template <int STAGE>
__device__ __forceinline__ void stage_update(const KParams &p, const KPtrs &q, uint32_t i, uint32_t w, float4 me,
                                             float fx, float fy, float fz, float4 in) {
  if (p.frozen) fx = fy = fz = 0.0f; // with no noise (relaxation) the bead stays put
  if (STAGE != 0 && p.fcap2 > 0.0f) {
    const float m2 = fx * fx + fy * fy + fz * fz;
    if (m2 > p.fcap2) {
      const float sc = sqrtf(p.fcap2 / m2);
      fx *= sc; fy *= sc; fz *= sc;
    }
  }
  if constexpr (STAGE == 0) {
    q.out0[i] = make_float4(fx, fy, fz, 0.0f);
  } else if constexpr (STAGE == 1) {
    const float nx = p.sd * in.x, ny = p.sd * in.y, nz = p.sd * in.z;
    const float ex = me.x + fx * p.dt + nx, ey = me.y + fy * p.dt + ny, ez = me.z + fz * p.dt + nz;
    q.out0[i] = make_float4(ex, ey, ez, me.w);
    q.pstar[i] = make_float4(me.x + fx * p.half_dt + nx, me.y + fy * p.half_dt + ny, me.z + fz * p.half_dt + nz,
                             0.0f);
    bin_bead(p, q, w, ex, ey, ez);
  } else {
    const float nx = in.x + fx * p.half_dt, ny = in.y + fy * p.half_dt, nz = in.z + fz * p.half_dt;
    q.out0[i] = make_float4(nx, ny, nz, me.w);
    bin_bead(p, q, w, nx, ny, nz);
  }
}

// stage_update with its input loaded here
template <int STAGE>
__device__ __forceinline__ void stage_epilogue(const KParams &p, const KPtrs &q, uint32_t i, uint32_t w,
                                               float4 me, float fx, float fy, float fz) {
  stage_update<STAGE>(p, q, i, w, me, fx, fy, fz, stage_input<STAGE>(p, q, i));
}

// ---------------------------------------------------------------- balanced force kernel
// Eight lanes per bead (LANES = 8). The bead's neighbour rows (2R+1 planes of
// 2R+1 rows) are trimmed to the cells that can hold a bead within its own
// cutoff (the het-het one for het beads, the short one otherwise); the slots
// of those cells are compacted into an LDS list, and the eight lanes take
// every eighth entry. A bead's candidates are then spread evenly over its
// lanes whatever the row lengths; three DPP steps sum the parts.
constexpr int C_LANES = 8;
// Workgroups of two wavefronts (16 beads): 256- and 512-lane ones group the
// types over more beads but schedule coarser, and measured 1.8 % and 9 %
// slower per step.
constexpr int BAL_BLOCK = 128;
constexpr int C_BEADS_PER_BLOCK = (64 / C_LANES) * (BAL_BLOCK / 64);
constexpr int C_LIST = 128; // candidate slots per bead held in LDS
static_assert(C_BEADS_PER_BLOCK <= 32, "the type grouping keeps one bit per bead in 32 bits");

// Squared list radius of bead w: its cutoff (the het-het one for het beads,
// the short one otherwise) plus 1e-3 sigma for the rounding of cell edges.
template <int MODE> __device__ __forceinline__ float trim_radius2(const KParams &p, uint32_t w) {
  if constexpr (MODE == MODE_ICG) {
    const float r = p.rco + 1e-3f;
    return r * r;
  }
  return (w & W_HET) ? p.trim2_long : p.trim2_short;
}

// Squared gaps along one axis between a bead at offset l from its cell's low
// edge and the cells k + 1 steps below (lo[k]) and above (hi[k]) its own.
// A cell outside the grid (c is the bead's cell on this axis) gets +inf, so
// it never passes a distance test and needs no bounds check of its own.
template <int R> struct AxisGaps {
  float lo[R], hi[R];
};
template <int R> __device__ __forceinline__ AxisGaps<R> axis_gaps(float l, float csl, int c, int cps) {
  AxisGaps<R> G;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const float a = fmaxf(l + (float)k * csl, 0.0f), b = fmaxf((float)(k + 1) * csl - l, 0.0f);
    G.lo[k] = c - (k + 1) >= 0 ? a * a : INFINITY;
    G.hi[k] = c + (k + 1) < cps ? b * b : INFINITY;
  }
  return G;
}
// The same squared gap for one offset d per lane: 0 for d = 0, the gap to
// the cell d steps away otherwise, +inf outside the grid.
__device__ __forceinline__ float axis_gap2(float l, float csl, int c, int d, int cps) {
  const float a = d < 0 ? l + (float)(-d - 1) * csl : (float)d * csl - l;
  const float g = d == 0 ? 0.0f : fmaxf(a, 0.0f);
  return (unsigned)(c + d) < (unsigned)cps ? g * g : INFINITY;
}
// v of lane k of the caller's eight-lane group (gb4: byte address of the
// group's lane 0 for ds_bpermute)
__device__ __forceinline__ float group_lane(float v, int gb4, int k) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(gb4 + (k << 2), __float_as_int(v)));
}

// Reach 2: the 25 neighbour rows (dy, dz) in ring order, the 3 x 3 inner
// rows first. A cutoff shorter than the cell side never reaches the outer
// ring, and a wavefront whose beads all have such cutoffs skips it.
constexpr int RING2[25][2] = {{0, 0},   {-1, 0}, {1, 0},  {0, -1}, {0, 1},  {-1, -1}, {1, -1},
                              {-1, 1},  {1, 1},  {-2, -2}, {-1, -2}, {0, -2}, {1, -2},  {2, -2},
                              {-2, -1}, {2, -1}, {-2, 0}, {2, 0},  {-2, 1}, {2, 1},   {-2, 2},
                              {-1, 2},  {0, 2},  {1, 2},  {2, 2}};
// The centre row (0, 0) is shared by the bead's eight lanes; the other 24
// go one per lane: rows 8 m + 1 .. 8 m + 8 of the ring order in word m, 6
// bits each, (dy + 2) | (dz + 2) << 3. Word 0 holds the inner ring.
constexpr uint64_t ring2_word(int m) {
  uint64_t w = 0;
  for (int k = 0; k < C_LANES; ++k) {
    const int r = m * C_LANES + k + 1;
    if (r < 25) w |= (uint64_t)((RING2[r][0] + 2) | ((RING2[r][1] + 2) << 3)) << (6 * k);
  }
  return w;
}
constexpr uint64_t RING2_WORDS[3] = {ring2_word(0), ring2_word(1), ring2_word(2)};

// Pair (and LBS) forces of the bead `me` over its trimmed candidate list,
// spread over the bead's G lanes (sub = 0..G-1; G = 8, or 2 for beads whose
// cutoff stays within one cell at reach 1) and added to f; on return every
// lane of the group holds the bead's total. `list` is the group's LDS list of
// LCAP entries, `self` the bead's own cell-order slot; the caller's lanes of
// one wavefront run this together. KIND: the beads' kind when the caller
// knows it (the two parts of the type-ordered kernel).
template <int MODE, bool BATCH, int R, int G = C_LANES, int LCAP = C_LIST, int KIND = BEAD_ANY>
__device__ __forceinline__ void bal_pair_sum(const KParams &p, const KPtrs &q, uint32_t *list, bool active,
                                             uint32_t self, int sub, int lane, float4 me, float &fx, float &fy,
                                             float &fz) {
  static_assert(G == 8 || (G == 2 && R == 1), "eight lanes per bead, or two at reach 1");
  static_assert(LCAP % (8 * G) == 0, "the gather rounds end on the list's last entry");
  // eight lanes: the group's lanes hold the y and z gaps of offsets -R..R
  // and each row fetches its two; two lanes compute all three themselves
  constexpr bool LANE_GAPS = G >= 2 * R + 1;
  // reach 2: rows in ring order, the outer ring skippable
  constexpr bool RING = R == 2;
  // NROW: the rows other than the centre one, RPL of them per lane
  constexpr int W = 2 * R + 1, NROW = W * W - 1, RPL = (NROW + G - 1) / G;
  const uint32_t w = wbits(me);
  const float rx = me.x, ry = me.y, rz = me.z;
  bool outside = false;
  const int cx = cell_coord(rx, p.gf, outside), cy = cell_coord(ry, p.gf, outside),
            cz = cell_coord(rz, p.gf, outside);
  const float rc2 = trim_radius2<MODE>(p, w);
  const float csl = p.gf.csl;
  const int cps = p.gf.cps;
  const AxisGaps<R> GX = axis_gaps<R>(rx - (float)(cx - p.gf.half) * csl, csl, cx, cps);
  // y and z: lane sub of the bead's group holds the squared gaps to the
  // cells d = sub - R steps away; each row takes its two from those lanes
  const int d = sub - R, gb4 = (lane & ~(G - 1)) << 2;
  float gyl = 0.0f, gzl = 0.0f;
  AxisGaps<R> GY, GZ;
  if constexpr (LANE_GAPS) {
    gyl = axis_gap2(ry - (float)(cy - p.gf.half) * csl, csl, cy, d, cps);
    gzl = axis_gap2(rz - (float)(cz - p.gf.half) * csl, csl, cz, d, cps);
  } else {
    GY = axis_gaps<R>(ry - (float)(cy - p.gf.half) * csl, csl, cy, cps);
    GZ = axis_gaps<R>(rz - (float)(cz - p.gf.half) * csl, csl, cz, cps);
  }
  // squared gap to the cells i - R steps away (i = 0 .. 2R) from the
  // bead's own gaps
  auto own_gap = [](const AxisGaps<R> &A, int i) {
    float g = 0.0f;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      g = i == R - 1 - k ? A.lo[k] : g;
      g = i == R + 1 + k ? A.hi[k] : g;
    }
    return g;
  };
  // cell (cx, cy + iy - R, cz + iz - R) is base0 + (iz cps + iy) cps; the
  // products take the 24-bit multiplier: grids are set up with cps^3 < 2^31,
  // so cps and iz cps + iy stay below 2^24
  const int base0 = p.gf.coff + ((cz - R) * cps + (cy - R)) * cps + cx;

  // Reach 2: the outer ring (rows of words 1 and 2) is skipped by a
  // wavefront in which no bead's cutoff reaches a cell two steps away in y
  // or z.
  bool any_outer = true;
  // (het beads always reach it: their cutoff spans two cells)
  if constexpr (RING && KIND != BEAD_HET) {
    const float m2 = fminf(gyl, gzl);
    const float near2 = fminf(group_lane(m2, gb4, 0), group_lane(m2, gb4, 2 * R));
    any_outer = __ballot(active && rc2 >= near2) != 0ull;
  }
  // The centre row (dy = dz = 0, both gaps 0, so h2 = rc2): the same cells
  // for the bead's eight lanes, listed first and shared by them.
  uint32_t oc0, oc1;
  {
    const int b = base0 + (R * cps + R) * cps;
    int c0 = b, c1 = b + 1;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      c0 -= GX.lo[k] <= rc2 ? 1 : 0;
      c1 += GX.hi[k] <= rc2 ? 1 : 0;
    }
    uint32_t a0 = active ? (uint32_t)c0 << 2 : 0u, a1 = active ? (uint32_t)c1 << 2 : 0u;
    asm("" : "+v"(a0), "+v"(a1));
    oc0 = a0;
    oc1 = a1;
  }
  // the other rows, sub, sub + 8, ... (reach 2: in ring order): the byte
  // offsets of their trimmed slot ranges first, then the start entries of
  // all of them loaded together, so that a lane waits once for all its rows
  uint32_t o0[RPL], o1[RPL];
#pragma unroll
  for (int m = 0; m < RPL; ++m) {
    o0[m] = 0u;
    o1[m] = 0u;
    if (RING && m >= 1 && !any_outer) continue;
    // Row (dy, dz) = (iy - R, iz - R): its cells [c0, c1) that can hold a
    // bead within sqrt(rc2), empty when the row's y-z slab is farther (or
    // outside the grid), else the run around cx whose x gaps fit in the rest,
    // rc2 - gap_y^2 - gap_z^2. Branch-free: an empty row selects [0, 0).
    const int r = sub + m * G;
    // row r of the W x W rows with the centre (index W * W / 2) left out
    const int rr = r < W * W / 2 ? r : r + 1;
    int iy = rr % W, iz = rr / W;
    if constexpr (RING) {
      const uint32_t v = (uint32_t)(RING2_WORDS[m] >> (6 * sub)) & 63u;
      iy = (int)(v & 7u);
      iz = (int)(v >> 3);
    }
    float h2;
    if constexpr (LANE_GAPS) {
      h2 = rc2 - group_lane(gyl, gb4, iy) - group_lane(gzl, gb4, iz);
    } else {
      h2 = rc2 - own_gap(GY, iy) - own_gap(GZ, iz);
    }
    const bool ok = active && r < NROW && h2 >= 0.0f;
    uint32_t t = __umul24((uint32_t)iz, (uint32_t)cps) + (uint32_t)iy;
    asm("" : "+v"(t));
    const int b = (int)__umul24(t, (uint32_t)cps) + base0;
    int c0 = b, c1 = b + 1;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      c0 -= GX.lo[k] <= h2 ? 1 : 0;
      c1 += GX.hi[k] <= h2 ? 1 : 0;
    }
    // hidden from the compiler so that the loads keep the scalar base and a
    // 32-bit offset register
    uint32_t a0 = ok ? (uint32_t)c0 << 2 : 0u, a1 = ok ? (uint32_t)c1 << 2 : 0u;
    asm("" : "+v"(a0), "+v"(a1));
    o0[m] = a0;
    o1[m] = a1;
  }
  uint32_t lo[RPL], len[RPL], tot = 0;
  const uint32_t clo = at_byte(q.fcstart, oc0);
  uint32_t clen = at_byte(q.fcstart, oc1);
#pragma unroll
  for (int m = 0; m < RPL; ++m) {
    lo[m] = at_byte(q.fcstart, o0[m]);
    len[m] = at_byte(q.fcstart, o1[m]);
  }
  clen -= clo;
#pragma unroll
  for (int m = 0; m < RPL; ++m) {
    len[m] -= lo[m];
    tot += len[m];
  }
  // the LBS sites of LAD beads, row by row (the centre row's on lane 0)
  if (!BATCH && MODE != MODE_ICG && p.n_l > 0 && (w & W_LAD)) {
    if (sub == 0) {
      const uint32_t llo = at_byte(q.lstart, oc0), lhi = at_byte(q.lstart, oc1);
      for (uint32_t k = llo; k < lhi; ++k) lbs_force(p, rx, ry, rz, q.lspos[k], fx, fy, fz);
    }
#pragma unroll
    for (int m = 0; m < RPL; ++m) {
      const uint32_t llo = at_byte(q.lstart, o0[m]), lhi = at_byte(q.lstart, o1[m]);
      for (uint32_t k = llo; k < lhi; ++k) lbs_force(p, rx, ry, rz, q.lspos[k], fx, fy, fz);
    }
  }
  // exclusive prefix of the row lengths over the bead's eight lanes
  // (row shifts by 1, 2, 4 within each 16-lane row, a lane taking only
  // values of its own group: the shift by 4 reaches lanes 4-7 of each group,
  // banks 1 and 3), and the group's total by the three DPP steps of the
  // force sum
  uint32_t incl = tot;
  {
    const uint32_t v1 = dpp_u<DPP_SHR1>(incl);
    incl += sub >= 1 ? v1 : 0u;
    if constexpr (G >= 4) {
      const uint32_t v2 = dpp_u<DPP_SHR2>(incl);
      incl += sub >= 2 ? v2 : 0u;
    }
    if constexpr (G == 8) incl += dpp_u<DPP_SHR4, 0xA>(incl);
  }
  uint32_t total = tot;
  total += dpp_u<DPP_XOR1>(total);
  if constexpr (G >= 4) total += dpp_u<DPP_XOR2>(total);
  if constexpr (G == 8) total += dpp_u<DPP_MIRROR8>(total);
  // the centre row's entries come first
  total += clen;
  uint32_t off = clen + incl - tot;
  const bool listed = total <= (uint32_t)LCAP;
  // The gathers run in rounds common to the wavefront, up to its longest
  // list: a wavefront issues every round that any of its groups needs, so
  // the shorter lists are padded with the bead's own slot (a term of exactly
  // zero: |i - j| <= 1 is excluded) instead of ending early. Each lane still
  // adds its own entries in list order, so a bead's sum does not depend on
  // the other beads of its wavefront.
  const uint32_t wmax = wave_max(listed ? total : 0u);
  const uint32_t s4 = self << 4;
  // the list holds byte offsets (slot x 16) into the ordered positions, so a
  // gather needs no address arithmetic
  if (listed) {
    for (uint32_t k = sub; k < clen; k += G) list[k] = (clo + k) << 4;
#pragma unroll
    for (int m = 0; m < RPL; ++m) {
      uint32_t b = lo[m] << 4;
      for (uint32_t k = 0; k < len[m]; ++k, b += 16u) list[off++] = b;
    }
    // pads up to the last entry a lane reads: wmax rounded up to G (LCAP is
    // a multiple of G, so they stay in the list)
    const uint32_t lim = (wmax + (uint32_t)G - 1u) / (uint32_t)G * (uint32_t)G;
    for (uint32_t k = total + sub; k < lim; k += G) list[k] = s4;
  }
  // the list is read by the lanes of the same wavefront only: complete the
  // LDS writes and order them before the reads, without a block barrier
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  const PairCtx pc = pair_ctx(p, me);
  if (listed) {
    // Rounds of eight gathers in flight per lane (entries n, n + G, n + 2G,
    // ...) while the wavefront's longest list holds 8G more, then one round
    // of just the gathers the rest needs (wmax and base are the same in all
    // lanes, so its width is a scalar branch per gather).
    uint32_t base = 0;
    for (; base + 8 * G <= wmax; base += 8 * G) {
      const uint32_t n = base + sub;
      float4 e[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) e[k] = at_byte(q.fspos, list[n + k * G]);
#pragma unroll
      for (int k = 0; k < 8; ++k) pair_term<MODE, KIND>(p, pc, e[k], fx, fy, fz);
    }
    const uint32_t rem = wmax - base, n = base + sub;
    if (rem > 0u) {
      // the eight list entries read together (all inside the group's LCAP
      // entries: base is a multiple of 8G below LCAP, a multiple of 8G), the
      // gathers only for those in use
      uint32_t j[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) j[k] = list[n + k * G];
      float4 e[8];
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if ((uint32_t)(k * G) < rem) e[k] = at_byte(q.fspos, j[k]);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if ((uint32_t)(k * G) < rem) pair_term<MODE, KIND>(p, pc, e[k], fx, fy, fz);
    }
  } else {
    // crowded neighbourhood: each lane scans its own rows, and every eighth
    // entry of the centre row
    for (uint32_t k = sub; k < clen; k += G) pair_term<MODE, KIND>(p, pc, q.fspos[clo + k], fx, fy, fz);
#pragma unroll
    for (int m = 0; m < RPL; ++m)
      for (uint32_t k = 0; k < len[m]; ++k) pair_term<MODE, KIND>(p, pc, q.fspos[lo[m] + k], fx, fy, fz);
  }
  // sum over the bead's eight lanes: quad swaps (lane ^ 1, lane ^ 2), then
  // the half-row mirror (lane 7 - i), which pairs each quad with the other;
  // every lane adds the same two values each time, so all end equal
  fx += dpp_f<DPP_XOR1>(fx); fy += dpp_f<DPP_XOR1>(fy); fz += dpp_f<DPP_XOR1>(fz);
  if constexpr (G >= 4) {
    fx += dpp_f<DPP_XOR2>(fx); fy += dpp_f<DPP_XOR2>(fy); fz += dpp_f<DPP_XOR2>(fz);
  }
  if constexpr (G == 8) {
    fx += dpp_f<DPP_MIRROR8>(fx); fy += dpp_f<DPP_MIRROR8>(fy); fz += dpp_f<DPP_MIRROR8>(fz);
  }
  // every lane's list reads are done before the group's next list is written
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ROUNDS = 1: each wavefront takes eight beads (eight lanes each); lane 0 of
// each bead adds bonds and wall and runs the stage (het beads first within a
// workgroup when p.group is set).
// ROUNDS > 1: each wavefront runs the pair phase for ROUNDS x 8 beads, eight
// at a time; lane 8 g + r keeps the total of bead g of round r, so the
// per-bead work after the pair sum (bonds, bending, wall, noise, update,
// binning) runs on ROUNDS x 8 lanes at once instead of 8.
// EPI = false (ROUNDS = 1 only): the kernel stops after the pair sum and
// writes it to q.fpair; k_integrate (one lane per bead) adds bonds and wall
// and runs the stage.
template <int MODE, int STAGE, bool BATCH, int R, bool EPI = true, int ROUNDS = 1>
__global__ __launch_bounds__(BAL_BLOCK) void k_force_bal(KParams p, KPtrs q) {
  static_assert(ROUNDS == 1 || EPI, "the split integration pass takes one round");
  static_assert(ROUNDS >= 1 && ROUNDS <= C_LANES, "a group keeps at most one bead per lane");
  __shared__ uint32_t s_list[C_BEADS_PER_BLOCK][C_LIST];
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  const int sub = lane % C_LANES, gw = lane / C_LANES, grp = wave * (64 / C_LANES) + gw;
  const int wbase = (xcd_block() * (BAL_BLOCK / 64) + wave) * ((64 / C_LANES) * ROUNDS);
  if constexpr (ROUNDS == 1) {
    int t = wbase + gw;
    float4 me = t < p.N ? q.fspos[t] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (p.group) {
      // A wavefront runs as long as its longest list. Het beads (cutoff 2.5
      // sigma) have about four times the candidates of the others, so the
      // workgroup's slots are reordered: het slots first, then the rest, each
      // in cell order. Position grp takes the k-th set bit of the het mask;
      // the positions move through LDS.
      __shared__ uint32_t s_mask[BAL_BLOCK / 64];
      __shared__ float4 s_me[C_BEADS_PER_BLOCK];
      const int t0 = wbase - wave * (64 / C_LANES);
      const bool het = t < p.N && (wbits(me) & W_HET) != 0u;
      const uint64_t b = __ballot(het && sub == 0);
      uint32_t m8 = 0;
#pragma unroll
      for (int k = 0; k < 64 / C_LANES; ++k) m8 |= (uint32_t)((b >> (k * C_LANES)) & 1u) << k;
      if (lane == 0) s_mask[wave] = m8;
      if (sub == 0) s_me[grp] = me;
      __syncthreads();
      uint32_t m = 0;
#pragma unroll
      for (int k = 0; k < BAL_BLOCK / 64; ++k) m |= s_mask[k] << (k * (64 / C_LANES));
      const int nh = __popc(m);
      uint32_t mm = grp < nh ? m : ~m;
      int k = grp < nh ? grp : grp - nh, pos = 0;
#pragma unroll
      for (int sh = C_BEADS_PER_BLOCK / 2; sh > 0; sh >>= 1) {
        const int c = __popc(mm & ((1u << sh) - 1u));
        if (k >= c) {
          k -= c;
          mm >>= sh;
          pos += sh;
        }
      }
      t = t0 + pos;
      me = s_me[pos];
    }
    const uint32_t w = wbits(me);
    const uint32_t i = w & ID_MASK;
    const bool active = t < p.N && i < (uint32_t)p.N; // a sentinel slot holds no bead
    if constexpr (BATCH) load_system(p, q, active ? (uint32_t)t : 0u);
    // the stage input is needed only at the end: load it now
    float4 in = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (EPI && active && sub == 0) in = stage_input<STAGE>(p, q, i);
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    if (EPI && active && sub == 0) {
      if (q.fbond) {
        const float4 fb = q.fbond[i];
        fx = fb.x; fy = fb.y; fz = fb.z;
      } else {
        bonded_force(p, q.pos, i, w, me.x, me.y, me.z, fx, fy, fz);
        wall_force<MODE>(p, me.x, me.y, me.z, w, fx, fy, fz);
      }
    }
    bal_pair_sum<MODE, BATCH, R>(p, q, s_list[grp], active, (uint32_t)t, sub, lane, me, fx, fy, fz);
    if (!active || sub != 0) return;
    if constexpr (!EPI) {
      q.fpair[i] = make_float4(fx, fy, fz, 0.0f);
    } else {
      stage_update<STAGE>(p, q, i, w, me, fx, fy, fz, in);
    }
  } else {
    float kx = 0.0f, ky = 0.0f, kz = 0.0f;
    for (int rd = 0; rd < ROUNDS; ++rd) {
      const int t = wbase + rd * (64 / C_LANES) + gw;
      const float4 me = t < p.N ? q.fspos[t] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      const bool active = t < p.N && (wbits(me) & ID_MASK) < (uint32_t)p.N;
      if constexpr (BATCH) load_system(p, q, active ? (uint32_t)t : 0u);
      float fx = 0.0f, fy = 0.0f, fz = 0.0f;
      bal_pair_sum<MODE, BATCH, R>(p, q, s_list[grp], active, (uint32_t)t, sub, lane, me, fx, fy, fz);
      if (sub == rd) {
        kx = fx;
        ky = fy;
        kz = fz;
      }
    }
    if (sub >= ROUNDS) return;
    const int t = wbase + sub * (64 / C_LANES) + gw;
    if (t >= p.N) return;
    const float4 me = q.fspos[t];
    const uint32_t w = wbits(me);
    const uint32_t i = w & ID_MASK;
    if (i >= (uint32_t)p.N) return;
    if constexpr (BATCH) load_system(p, q, (uint32_t)t);
    if (q.fbond) {
      const float4 fb = q.fbond[i];
      kx += fb.x; ky += fb.y; kz += fb.z;
    } else {
      bonded_force(p, q.pos, i, w, me.x, me.y, me.z, kx, ky, kz);
      wall_force<MODE>(p, me.x, me.y, me.z, w, kx, ky, kz);
    }
    stage_epilogue<STAGE>(p, q, i, w, me, kx, ky, kz);
  }
}

// The balanced kernel over a type-ordered work list (q.work: the cell-order
// slots of the het beads, then those of the others, each part in cell order,
// written by the order pass). Every wavefront holds beads of one type:
// - the first p.het_blocks workgroups take the het beads, eight lanes per
//   bead over the 5 x 5 x 5 half-size cells (their cutoff spans two cells);
// - the rest take the others, two lanes per bead over 3 x 3 x 3 of the same
//   cells (their cutoff is shorter than one cell side, checked by the host),
//   64 beads per workgroup with lists of up to T_LIST entries. Their lists
//   hold ~19 candidates, so two lanes keep the gathers in rounds while each
//   bead's set-up (rows, list, epilogue) is repeated on two lanes, not four
//   or eight.
// Each part's workgroups are renumbered per XCD on their own, so every XCD
// gets its share of both. After the pair sums, the first wavefront of a
// workgroup runs the stage update and binning of both wavefronts' beads.
// Results are those of k_force_bal: the same pair set, summed in another
// order.
constexpr int H_LANES = C_LANES;                       // lanes per bead of the het part
constexpr int H_BEADS_PER_BLOCK = BAL_BLOCK / H_LANES; // 16
constexpr int T_LANES = 2;                             // lanes per bead of the non-het part
constexpr int T_BEADS_PER_BLOCK = BAL_BLOCK / T_LANES; // 64
constexpr int T_LIST = 48;                             // list entries per bead of the non-het part
constexpr int T_LDS = H_BEADS_PER_BLOCK * C_LIST > T_BEADS_PER_BLOCK * T_LIST ? H_BEADS_PER_BLOCK * C_LIST
                                                                              : T_BEADS_PER_BLOCK * T_LIST;
static_assert(H_BEADS_PER_BLOCK == TYPED_HET_BEADS_PER_BLOCK, "het workgroups as the host sizes them");
static_assert(BAL_BLOCK == 128, "the epilogue hand-over pairs the workgroup's two wavefronts");
static_assert(8 * C_LIST * 4 >= 8 * 3 * 16 && 32 * T_LIST * 4 >= 32 * 3 * 16,
              "a wavefront's list memory holds the hand-over records of its beads");
template <int MODE, int STAGE, bool BATCH>
__global__ __launch_bounds__(BAL_BLOCK) void k_force_typed(KParams p, KPtrs q, NoiseJob nj) {
  __shared__ __align__(16) uint32_t s_list[T_LDS];
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  const int b = (int)blockIdx.x;
  const int t_blocks = (p.N - p.n_het + T_BEADS_PER_BLOCK - 1) / T_BEADS_PER_BLOCK;
  // Workgroups past the force ones (when the launch has them) draw the next
  // step's normals, one lane per bead: dispatched last, they run in the
  // launch's tail, on SIMDs that the force wavefronts leave idle.
  if (b >= p.het_blocks + t_blocks) {
    const uint32_t g = (uint32_t)(b - p.het_blocks - t_blocks) * BAL_BLOCK + threadIdx.x;
    if (nj.z && g < nj.n) fill_noise(nj, g);
    return;
  }
  const bool het_part = b < p.het_blocks;
  int t, sub, grp;
  uint32_t end;
  if (het_part) {
    sub = lane % H_LANES;
    grp = wave * (64 / H_LANES) + lane / H_LANES;
    t = xcd_remap(b, p.het_blocks) * H_BEADS_PER_BLOCK + grp;
    end = (uint32_t)p.n_het;
  } else {
    sub = lane % T_LANES;
    grp = wave * (64 / T_LANES) + lane / T_LANES;
    t = p.n_het + xcd_remap(b - p.het_blocks, t_blocks) * T_BEADS_PER_BLOCK + grp;
    end = (uint32_t)p.N;
  }
  const bool in = (uint32_t)t < end;
  // the work list holds slots below N; min() keeps a stale entry (left by an
  // overflowing cell, which is reported) inside the arrays
  const uint32_t slot = in ? min(q.work[t], (uint32_t)p.N - 1u) : 0u;
  const float4 me = in ? q.fspos[slot] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const uint32_t w = wbits(me);
  const uint32_t i = w & ID_MASK;
  const bool active = in && i < (uint32_t)p.N; // a sentinel slot holds no bead
  if constexpr (BATCH) load_system(p, q, active ? slot : 0u);
  const bool lead = sub == 0;
  float4 inp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (active && lead) inp = stage_input<STAGE>(p, q, i);
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  if (active && lead) {
    if (q.fbond) {
      const float4 fb = q.fbond[i];
      fx = fb.x; fy = fb.y; fz = fb.z;
    } else {
      bonded_force(p, q.pos, i, w, me.x, me.y, me.z, fx, fy, fz);
      wall_force<MODE>(p, me.x, me.y, me.z, w, fx, fy, fz);
    }
  }
  if (het_part)
    bal_pair_sum<MODE, BATCH, 2, H_LANES, C_LIST, BEAD_HET>(p, q, s_list + grp * C_LIST, active, slot, sub, lane, me,
                                                             fx, fy, fz);
  else
    bal_pair_sum<MODE, BATCH, 1, T_LANES, T_LIST, BEAD_SHORT>(p, q, s_list + grp * T_LIST, active, slot, sub, lane,
                                                              me, fx, fy, fz);
  // The epilogue of both wavefronts' beads runs on the first one, so the
  // per-bead work is issued once per workgroup instead of once per
  // wavefront: the second hands over each bead's total, slot, position and
  // stage input through its own (finished) part of the list memory, and
  // lane 1 of each group of the first takes the bead of the same group of
  // the second.
  const int gw = lane / (het_part ? H_LANES : T_LANES);
  float4 *xch = reinterpret_cast<float4 *>(s_list + (het_part ? 8 * C_LIST : 32 * T_LIST));
  if (wave == 1 && lead) {
    xch[3 * gw] = make_float4(fx, fy, fz, __uint_as_float(active ? slot : 0xffffffffu));
    xch[3 * gw + 1] = me;
    xch[3 * gw + 2] = inp;
  }
  __syncthreads();
  if (wave == 1) return;
  bool run = active && lead;
  float4 mm = me, in2 = inp;
  if (sub == 1) {
    const float4 F = xch[3 * gw];
    const uint32_t s2 = __float_as_uint(F.w);
    run = s2 != 0xffffffffu;
    fx = F.x;
    fy = F.y;
    fz = F.z;
    mm = xch[3 * gw + 1];
    in2 = xch[3 * gw + 2];
    if constexpr (BATCH) load_system(p, q, run ? s2 : 0u);
  }
  if (!run) return;
  const uint32_t w2 = wbits(mm);
  stage_update<STAGE>(p, q, w2 & ID_MASK, w2, mm, fx, fy, fz, in2);
}

// Bonds, wall and the Heun stage of every bead, one lane per bead in index
// order, on top of the pair forces k_force_bal<..., false> left in q.fpair.
template <int MODE, int STAGE, bool BATCH>
__global__ __launch_bounds__(BEAD_BLOCK) void k_integrate(KParams p, KPtrs q) {
  const int t = xcd_block() * BEAD_BLOCK + threadIdx.x;
  if (t >= p.N) return;
  const uint32_t i = (uint32_t)t;
  const float4 me = q.pos[i];
  const uint32_t w = wbits(me);
  if constexpr (BATCH) load_system(p, q, i);
  const float4 fp = q.fpair[i];
  float fx = fp.x, fy = fp.y, fz = fp.z;
  bonded_force(p, q.pos, i, w, me.x, me.y, me.z, fx, fy, fz);
  wall_force<MODE>(p, me.x, me.y, me.z, w, fx, fy, fz);
  stage_epilogue<STAGE>(p, q, i, w, me, fx, fy, fz);
}

// Binning of an index-ordered array (initial builds, LBS grid, export grid).
__global__ __launch_bounds__(BEAD_BLOCK) void k_bin(GridGeom g, int n_obj, const float4 *pos, int het_only,
                                                    uint32_t *count, uint32_t *blocksum, float4 *bucket,
                                                    uint32_t *flags, uint32_t *hblocksum) {
  const int i = blockIdx.x * BEAD_BLOCK + threadIdx.x;
  if (i >= n_obj) return;
  const float4 v = pos[i];
  if (het_only && !(wbits(v) & W_HET)) return;
  bin_into(g, count, blocksum, bucket, flags, v, hblocksum);
}

// Binning of the index-ordered beads of a batch into their systems' grids
// (bead i belongs to the system whose slot range holds i).
__global__ __launch_bounds__(BEAD_BLOCK) void k_bin_batch(const SysParams *sys, const uint32_t *sys_slot, int n_sys,
                                                          int n, const float4 *pos, uint32_t *count,
                                                          uint32_t *blocksum, float4 *bucket, uint32_t *flags,
                                                          uint32_t *hblocksum) {
  const int i = blockIdx.x * BEAD_BLOCK + threadIdx.x;
  if (i >= n) return;
  const float4 v = pos[i];
  const SysParams &S = sys[system_of(sys_slot, n_sys, (uint32_t)i)];
  bin_into(S.g, count, blocksum, bucket, flags, v, hblocksum);
}

// Per-cell scan + deterministic order of one or two grids (replaces ccp's
// calc_indexes/SortPairs/set_cells_empty/find_cells_limits sequence,
// sugrid.cu:125-133). The bucket holds each member's position with its index
// in the w bits. The block's entries are staged in LDS; each entry's rank
// within its cell (number of smaller indices) is then found in parallel and
// the entry written to its sorted slot.
constexpr int ORDER_LDS = 1024;

__global__ __launch_bounds__(SCAN_BLOCK) void k_order(OrderJob ja, OrderJob jb, uint32_t *flags,
                                                      uint64_t *step_inc, NoiseJob nj, KParams bp, BondJob bj) {
  __shared__ uint32_t red[SCAN_BLOCK / 64], wsum[SCAN_BLOCK / 64], msum[SCAN_BLOCK / 64];
  __shared__ uint32_t hred[SCAN_BLOCK / 64], hsum[SCAN_BLOCK / 64];
  __shared__ uint32_t s_id[ORDER_LDS];
  __shared__ float4 s_pos[ORDER_LDS];
  __shared__ uint8_t s_cell[ORDER_LDS];
  __shared__ uint32_t s_loff[SCAN_BLOCK + 1];
  __shared__ uint32_t s_start[SCAN_BLOCK];
  __shared__ uint32_t s_hoff[SCAN_BLOCK];
  const int bx = xcd_block();
  const bool second = bx >= ja.g.nblk;
  const OrderJob &J = second ? jb : ja;
  const int b = second ? bx - ja.g.nblk : bx;
  const int tid = threadIdx.x, lane = __lane_id(), wid = tid >> 6;
  const int c = b * SCAN_BLOCK + tid;
  const int cap = J.g.cap, n_c = J.g.n_c;
  // type-ordered work list of this grid (workgroup-uniform)
  const bool typed = J.work != nullptr;
  // The ordering's first loads go out before the per-bead terms below, so
  // that their latency overlaps that work: this cell's count and the totals
  // of the workgroups before this one (four per lane at once; with the work
  // list also their het totals).
  constexpr int PRE = 4;
  const uint32_t cnt = (c < n_c) ? J.count[c] : 0u;
  uint32_t pv[PRE], hv[PRE];
#pragma unroll
  for (int u = 0; u < PRE; ++u) {
    const int k = tid + u * SCAN_BLOCK;
    pv[u] = k < b ? J.bs_read[k] : 0u;
    hv[u] = typed && k < b ? J.hbs_read[k] : 0u;
  }
  // Per-bead terms of the next force pass at one lane per bead, where the
  // force kernels would run them on one lane of eight: the next step's
  // normals (nj.z) and the bonded + wall forces of the positions just
  // ordered (bj.out). The first N / 256 workgroups take them, in bead order,
  // ahead of the ordering so that their arithmetic covers its first loads
  // (workgroups of their own beside the ordering ones measured slower).
  // A bead's positions for its bonded terms are loaded before its normals
  // are drawn, so that the two overlap.
  {
    const uint32_t n_terms = max(nj.z ? nj.n : 0u, bj.out ? bj.n : 0u);
    for (uint32_t g = blockIdx.x * SCAN_BLOCK + tid; g < n_terms; g += gridDim.x * SCAN_BLOCK) {
      const bool bond = bj.out && g < bj.n;
      BondNbrs nbr;
      if (bond) nbr = load_bond_nbrs(bj, g);
      if (nj.z && g < nj.n) fill_noise(nj, g);
      if (bond) fill_bond(bp, bj, g, nbr);
    }
  }
  uint32_t part = (pv[0] + pv[1]) + (pv[2] + pv[3]);
  uint32_t hpart = (hv[0] + hv[1]) + (hv[2] + hv[3]);
  for (int k0 = tid + PRE * SCAN_BLOCK; k0 < b; k0 += PRE * SCAN_BLOCK) {
#pragma unroll
    for (int u = 0; u < PRE; ++u) {
      const int k = k0 + u * SCAN_BLOCK;
      pv[u] = k < b ? J.bs_read[k] : 0u;
      hv[u] = typed && k < b ? J.hbs_read[k] : 0u;
    }
    part += (pv[0] + pv[1]) + (pv[2] + pv[3]);
    hpart += (hv[0] + hv[1]) + (hv[2] + hv[3]);
  }
  part = wave_sum(part);
  if (lane == 0) red[wid] = part;
  if (typed) {
    hpart = wave_sum(hpart);
    if (lane == 0) hred[wid] = hpart;
  }
  uint32_t m = cnt;
  if (m > (uint32_t)cap) {
    atomicAdd(flags + FLAG_OVERFLOW, 1u);
    m = (uint32_t)cap;
  }
  // the cell's first entries, loaded together now so that their latency
  // overlaps the scans and the barrier below
  const float4 *bk = J.bucket + (size_t)c * cap;
  constexpr uint32_t EARLY = 4;
  float4 ev[EARLY];
#pragma unroll
  for (uint32_t k = 0; k < EARLY; ++k)
    if (k < m) ev[k] = bk[k];
  const uint32_t incl = wave_incl_scan(cnt);
  const uint32_t mincl = wave_incl_scan(m);
  if (lane == 63) {
    wsum[wid] = incl;
    msum[wid] = mincl;
  }
  // het members of this cell (the work list places them before the others)
  uint32_t hc = 0, hincl = 0;
  if (typed) {
#pragma unroll
    for (uint32_t k = 0; k < EARLY; ++k)
      if (k < m) hc += (wbits(ev[k]) & W_HET) ? 1u : 0u;
    for (uint32_t k = EARLY; k < m; ++k) hc += (wbits(bk[k]) & W_HET) ? 1u : 0u;
    hincl = wave_incl_scan(hc);
    if (lane == 63) hsum[wid] = hincl;
  }
  __syncthreads();
  uint32_t prefix = 0, woff = 0, moff = 0, mtotal = 0, hprefix = 0, hoff = 0;
#pragma unroll
  for (int k = 0; k < SCAN_BLOCK / 64; ++k) {
    prefix += red[k];
    mtotal += msum[k];
    if (k < wid) {
      woff += wsum[k];
      moff += msum[k];
    }
  }
  if (typed) {
#pragma unroll
    for (int k = 0; k < SCAN_BLOCK / 64; ++k) {
      hprefix += hred[k];
      if (k < wid) hoff += hsum[k];
    }
    // het slots before this cell's: those of the blocks before and of the
    // cells before it in this block
    hoff += hprefix + hincl - hc;
  }
  const uint32_t start = prefix + woff + incl - cnt;
  moff += mincl - m;
  if (c < n_c) {
    J.cstart[c] = start;
    J.count[c] = 0u;
  }
  if (c == n_c - 1) J.cstart[n_c] = start + cnt;
  if (tid == 0) {
    J.bs_clear[b] = 0u;
    // the grid's het totals are cleared by every ordering of its buffers
    // (those with no work list too), so that the next binning finds them zero
    if (J.hbs_clear) J.hbs_clear[b] = 0u;
    if (b == 0 && !second && step_inc) *step_inc += 1ull;
  }
  // work-list entry of the member at slot dst, with hbefore het members in
  // the slots before it: het members in slot order, then the others
  auto put_work = [&](uint32_t dst, uint32_t hbefore, bool het) {
    const uint32_t k = het ? hbefore : J.n_het + (dst - hbefore);
    if (k < J.n_slots) J.work[k] = dst;
  };
  if (mtotal <= (uint32_t)ORDER_LDS) {
    s_loff[tid] = moff;
    s_start[tid] = start;
    s_hoff[tid] = hoff;
    if (tid == SCAN_BLOCK - 1) s_loff[SCAN_BLOCK] = moff + m;
    auto put = [&](uint32_t k, float4 v) {
      s_pos[moff + k] = v;
      s_id[moff + k] = __float_as_uint(v.w) & ID_MASK;
      s_cell[moff + k] = (uint8_t)tid;
    };
#pragma unroll
    for (uint32_t k = 0; k < EARLY; ++k)
      if (k < m) put(k, ev[k]);
    for (uint32_t k = EARLY; k < m; ++k) put(k, bk[k]);
    __syncthreads();
    for (uint32_t e = tid; e < mtotal; e += SCAN_BLOCK) {
      const uint32_t l = s_cell[e], id = s_id[e];
      const uint32_t b0 = s_loff[l], b1 = s_loff[l + 1];
      uint32_t rank = 0, hr = 0;
      if (typed) {
        for (uint32_t s2 = b0; s2 < b1; ++s2) {
          const bool before = s_id[s2] < id;
          rank += before ? 1u : 0u;
          hr += (before && (wbits(s_pos[s2]) & W_HET)) ? 1u : 0u;
        }
      } else {
        for (uint32_t s2 = b0; s2 < b1; ++s2) rank += (s_id[s2] < id) ? 1u : 0u;
      }
      const uint32_t dst = s_start[l] + rank;
      const float4 v = s_pos[e];
      J.spos[dst] = v;
      if (J.soi) J.soi[dst] = id;
      if (typed) put_work(dst, s_hoff[l] + hr, (wbits(v) & W_HET) != 0u);
    }
  } else {
    // dense block (raw initial walks): rank each cell's entries in place
    for (uint32_t k = 0; k < m; ++k) {
      const float4 v = bk[k];
      const uint32_t e = __float_as_uint(v.w) & ID_MASK;
      uint32_t rank = 0, hr = 0;
      for (uint32_t s2 = 0; s2 < m; ++s2) {
        const float4 o = bk[s2];
        const bool before = (__float_as_uint(o.w) & ID_MASK) < e;
        rank += before ? 1u : 0u;
        hr += (before && (wbits(o) & W_HET)) ? 1u : 0u;
      }
      J.spos[start + rank] = v;
      if (J.soi) J.soi[start + rank] = e;
      if (typed) put_work(start + rank, hoff + hr, (wbits(v) & W_HET) != 0u);
    }
  }
  // A cell with more members than its bucket holds (flagged above; the host
  // then grows the buckets and rebuilds) keeps its full slot range: the slots
  // past the stored members get a sentinel that every reader skips (index
  // ID_MASK, far outside any grid) instead of stale memory.
  for (uint32_t k = m; k < cnt; ++k) {
    J.spos[start + k] = make_float4(1e18f, 1e18f, 1e18f, __uint_as_float(ID_MASK));
    if (J.soi) J.soi[start + k] = ID_MASK;
    if (typed) put_work(start + k, hoff + hc, false);
  }
}

// The same terms on their own, before the first stage of a run.
__global__ __launch_bounds__(BEAD_BLOCK) void k_bead_terms(NoiseJob nj, KParams p, BondJob bj) {
  const uint32_t g = blockIdx.x * BEAD_BLOCK + threadIdx.x;
  if (nj.z && g < nj.n) fill_noise(nj, g);
  if (bj.out && g < bj.n) fill_bond(p, bj, g, load_bond_nbrs(bj, g));
}

// Contact counts on the fine grid: pairs i < j with d < rc within
// rad = ceil(rc / csl) cells, FP32 without fused operations so the decision
// matches the oracle bit for bit. Bead i (index within its system, whose
// first bead is id0) counts its partners j > i into bins of bin_size beads.
__device__ __forceinline__ void bead_contacts(const GridGeom &g, const float4 *spos, const uint32_t *cstart,
                                              float4 me, uint32_t i, uint32_t id0, int rad, float rc, uint32_t nb,
                                              uint32_t bin_size, uint32_t *map) {
#pragma clang fp contract(off)
  bool outside = false;
  const int cx = cell_coord(me.x, g, outside), cy = cell_coord(me.y, g, outside), cz = cell_coord(me.z, g, outside);
  const uint32_t bi = i / bin_size;
  for (int dz = -rad; dz <= rad; ++dz) {
    const int z = cz + dz;
    if (z < 0 || z >= g.cps) continue;
    for (int dy = -rad; dy <= rad; ++dy) {
      const int y = cy + dy;
      if (y < 0 || y >= g.cps) continue;
      const int row = g.coff + (z * g.cps + y) * g.cps;
      const uint32_t lo = cstart[row + max(cx - rad, 0)], hi = cstart[row + min(cx + rad, g.cps - 1) + 1];
      for (uint32_t k = lo; k < hi; ++k) {
        const float4 o = spos[k];
        const uint32_t j = (wbits(o) & ID_MASK) - id0;
        if (j <= i) continue;
        const float dx = me.x - o.x, dy2 = me.y - o.y, dz2 = me.z - o.z;
        const float xx = dx * dx, yy = dy2 * dy2, zz = dz2 * dz2;
        const float d2 = (xx + yy) + zz;
        if (sqrtf(d2) < rc) {
          const uint32_t bj = j / bin_size;
          if (bi < nb && bj < nb) {
            atomicAdd(map + (size_t)bi * nb + bj, 1u);
            atomicAdd(map + (size_t)bj * nb + bi, 1u);
          }
        }
      }
    }
  }
}

__global__ __launch_bounds__(BEAD_BLOCK) void k_contacts(KParams p, const float4 *spos, const uint32_t *cstart,
                                                         uint32_t *map, int rad, uint32_t slot0, uint32_t id0) {
  const int t = xcd_block() * BEAD_BLOCK + threadIdx.x;
  if (t >= p.N) return;
  const float4 me = spos[slot0 + t];
  bead_contacts(p.gf, spos, cstart, me, (wbits(me) & ID_MASK) - id0, id0, rad, p.contact_rc,
                (uint32_t)p.contact_bins, (uint32_t)p.contact_bin_size, map);
}

// The contacts of every system of a batch in one launch: slot t belongs to
// the system whose slot range holds it; that system's map is maps + s bins^2
// and its bins hold ceil(n / bins) beads.
__global__ __launch_bounds__(BEAD_BLOCK) void k_contacts_batch(const SysParams *sys, const uint32_t *sys_slot,
                                                               int n_sys, int n, float rc, uint32_t bins,
                                                               const float4 *spos, const uint32_t *cstart,
                                                               uint32_t *maps) {
  const int t = xcd_block() * BEAD_BLOCK + threadIdx.x;
  if (t >= n) return;
  const int s = system_of(sys_slot, n_sys, (uint32_t)t);
  const SysParams &S = sys[s];
  const float4 me = spos[t];
  const int rad = (int)ceilf(rc / S.g.csl);
  bead_contacts(S.g, spos, cstart, me, (wbits(me) & ID_MASK) - S.slot, S.slot, rad, rc, bins,
                (S.n + bins - 1) / bins, maps + (size_t)s * bins * bins);
}

// Non-bonded pairs (|i - j| > 1) closer than dmin <= fine cell side
// (particle_overlaps, chrsim.cu:805-820), each pair counted once.
__global__ __launch_bounds__(BEAD_BLOCK) void k_overlaps(KParams p, const float4 *spos, const uint32_t *cstart,
                                                         uint32_t slot0, float dmin2, unsigned long long *count) {
  const int t = blockIdx.x * BEAD_BLOCK + threadIdx.x;
  uint32_t n = 0;
  if (t < p.N) {
    const float4 me = spos[slot0 + t];
    const uint32_t i = wbits(me) & ID_MASK;
    Rows R;
    neighbour_rows(p.gf, cstart, me.x, me.y, me.z, R);
    for (int r = 0; r < 9; ++r)
      for (uint32_t k = R.lo[r]; k < R.hi[r]; ++k) {
        const float4 o = spos[k];
        const uint32_t j = wbits(o) & ID_MASK;
        const float dx = me.x - o.x, dy = me.y - o.y, dz = me.z - o.z;
        if (j > i + 1u && dist2(dx, dy, dz) < dmin2) ++n;
      }
  }
  const uint32_t tot = wave_sum(n);
  if (__lane_id() == 0 && tot) atomicAdd(count, (unsigned long long)tot);
}

__global__ __launch_bounds__(BEAD_BLOCK) void k_noise(KParams p, uint64_t step, float4 *out) {
  const int i = blockIdx.x * BEAD_BLOCK + threadIdx.x;
  if (i >= p.N) return;
  float z0, z1, z2;
  const uint32_t a = noise3(p.seed_lo, p.seed_hi, (uint32_t)i, step, z0, z1, z2);
  out[i] = make_float4(z0, z1, z2, __uint_as_float(a));
}

// ccp's per-frame observables (chrana.cu:372-480) summed in FP64: out[0..2]
// sum of r, [3] sum of |r|^2, [4] sum of the bond order P2(cos), [5] beads
// with z < nod, then the radial density of A, B and all beads (3 x n_b,
// lens volume of a bead of radius a inside the sphere of each bin edge,
// differenced bin by bin).
// Volume of a bead of radius a at distance d from the centre inside each of
// n_b shells of equal volume (edges R_n cbrt((k + 1) / n_b)), as lens
// volumes at the outer edge differenced bin by bin; add(k, net) receives the
// nonzero differences.
template <class F> __device__ __forceinline__ void lens_bins(double d, double a, double R_n, int n_b, F &&add) {
  const double vb = 4.0 / 3.0 * M_PI * a * a * a;
  double prev = 0.0;
  for (int k = 0; k < n_b; ++k) {
    const double rb = R_n * cbrt((k + 1.0) / n_b);
    double v;
    if (d - a > rb) v = 0.0;
    else if (d + a < rb) v = vb;
    else {
      const double ca = (d * d + a * a - rb * rb) / (2.0 * d * a);
      const double cb = (rb * rb + d * d - a * a) / (2.0 * rb * d);
      v = M_PI / 3.0 * a * a * a * (2.0 + ca) * (1.0 - ca) * (1.0 - ca) +
          M_PI / 3.0 * rb * rb * rb * (2.0 + cb) * (1.0 - cb) * (1.0 - cb);
    }
    const double net = v - prev;
    prev = v;
    if (net != 0.0) add(k, net);
    if (v == vb) break; // the rest of the bins see no change
  }
}

__global__ __launch_bounds__(BEAD_BLOCK) void k_observables(int N, double R_n, double nod, double a, int n_b,
                                                            const float4 *pos, double *out) {
  extern __shared__ double s_acc[];
  const int tid = threadIdx.x, n_acc = 6 + 3 * n_b;
  for (int k = tid; k < n_acc; k += BEAD_BLOCK) s_acc[k] = 0.0;
  __syncthreads();
  const int i = blockIdx.x * BEAD_BLOCK + tid;
  if (i < N) {
    const float4 q = pos[i];
    const double x = q.x, y = q.y, z = q.z;
    atomicAdd(&s_acc[0], x);
    atomicAdd(&s_acc[1], y);
    atomicAdd(&s_acc[2], z);
    const double d2 = x * x + y * y + z * z;
    atomicAdd(&s_acc[3], d2);
    if (i + 1 < N) {
      const float4 n = pos[i + 1];
      const double vx = (double)n.x - x, vy = (double)n.y - y, vz = (double)n.z - z;
      const double c = (vx * x + vy * y + vz * z) / (sqrt(vx * vx + vy * vy + vz * vz) * sqrt(d2));
      atomicAdd(&s_acc[4], 0.5 * (3.0 * c * c - 1.0));
    }
    if (z < nod) atomicAdd(&s_acc[5], 1.0);
    const int t = (wbits(q) & W_HET) ? 0 : 1;
    lens_bins(sqrt(d2), a, R_n, n_b, [&](int k, double net) {
      atomicAdd(&s_acc[6 + t * n_b + k], net);
      atomicAdd(&s_acc[6 + 2 * n_b + k], net);
    });
  }
  __syncthreads();
  for (int k = tid; k < n_acc; k += BEAD_BLOCK)
    if (s_acc[k] != 0.0) atomicAdd(&out[k], s_acc[k]);
}

// Radial density of every system of a batch (the all-bead row of
// k_observables), one lane per bead, FP64 sums into out[s n_b + k].
__global__ __launch_bounds__(BEAD_BLOCK) void k_radial_batch(const SysParams *sys, const uint32_t *sys_slot,
                                                             int n_sys, int n, double a, int n_b, const float4 *pos,
                                                             double *out) {
  const int i = blockIdx.x * BEAD_BLOCK + threadIdx.x;
  if (i >= n) return;
  const int s = system_of(sys_slot, n_sys, (uint32_t)i);
  const float4 q = pos[i];
  const double x = q.x, y = q.y, z = q.z;
  double *o = out + (size_t)s * n_b;
  lens_bins(sqrt(x * x + y * y + z * z), a, (double)sys[s].R_n, n_b, [&](int k, double net) { atomicAdd(o + k, net); });
}

// ccp's d(s) and 1000 P(s) of one chromosome (calc_sd_cp, chrana.cu:13-45):
// one thread per separation s = ia + 1, FP32 sums over the pairs (p, p+s) of
// beads [i_s, i_e) and, when N = 30386, of the homologue N/2 further on.
// Writes the frame's values to sd/cp and, when given, adds them to the FP64
// per-simulation sums as ccp's host loop does (chrana.cu:496-502).
__global__ __launch_bounds__(BEAD_BLOCK) void k_sd_cp(const float4 *pos, uint32_t N, uint32_t lsdcp, uint32_t i_s,
                                                      uint32_t i_e, float *sd, float *cp, double *sum_sd,
                                                      double *sum_cp) {
#pragma clang fp contract(off)
  const uint32_t ia = blockIdx.x * BEAD_BLOCK + threadIdx.x;
  if (ia >= lsdcp) return;
  const uint32_t s = ia + 1, ncc = (N == 30386u) ? 2u : 1u;
  float a = 0.0f, c = 0.0f;
  for (uint32_t k = 0; k < ncc; ++k) {
    for (uint32_t p = i_s; p < i_e - s; ++p) {
      const float4 u = pos[p], v = pos[p + s];
      const float dx = v.x - u.x, dy = v.y - u.y, dz = v.z - u.z;
      const float xx = dx * dx, yy = dy * dy, zz = dz * dz;
      const float d = sqrtf((xx + yy) + zz);
      a += d;
      if (d < 2.0f) c += 1000.0f;
    }
    i_s += N / ncc;
    i_e += N / ncc;
  }
  const float m = (float)(ncc * (i_e - i_s - s));
  a /= m;
  c /= m;
  if (sd) {
    sd[ia] = a;
    cp[ia] = c;
  }
  if (sum_sd) {
    sum_sd[ia] += (double)a;
    sum_cp[ia] += (double)c;
  }
}

// ccp's per-frame contact map (calc_cm, chrana.cu:48-84) with its exact
// float arithmetic: pixel ia of the packed lower triangle, 4-bead pixels,
// contact if |r_p - r_q| < 2, factor 1000, average over the two chromosome
// copies when N = 30386. Writes the frame's map to cm and, when given, adds
// it to the FP64 per-simulation sum (chrana.cu:505-511).
__global__ __launch_bounds__(BEAD_BLOCK) void k_ccp_cm(uint32_t N, uint32_t lcm, const float4 *pos, float *cm,
                                                       double *sum) {
#pragma clang fp contract(off)
  const uint32_t ia = blockIdx.x * BEAD_BLOCK + threadIdx.x;
  if (ia >= lcm) return;
  const uint32_t px = 4, ncc = (N == 30386u) ? 2u : 1u, cms = (N / ncc) / px;
  // ccp's pixel coordinates; the clamps keep the reads inside the chain
  // where the FP32 root of a large index would round past the last row
  const uint32_t ixp = min((uint32_t)((double)sqrtf((float)(2.0 * ia + 0.25)) - 0.5), cms - 1);
  const uint32_t iyp = min((uint32_t)(ia - ixp * (ixp + 1.0) / 2.0), cms - 1);
  uint32_t sx = ixp * px, ex = sx + px, sy = iyp * px, ey = sy + px;
  float acc = 0.0f;
  for (uint32_t c = 0; c < ncc; ++c) {
    for (uint32_t p = sx; p < ex; ++p)
      for (uint32_t q = sy; q < ey; ++q) {
        const float4 u = pos[p], v = pos[q];
        const float dx = u.x - v.x, dy = u.y - v.y, dz = u.z - v.z;
        const float xx = dx * dx, yy = dy * dy, zz = dz * dz;
        const float d = sqrtf((xx + yy) + zz);
        if (d < 2.0f) acc += 1000.0f;
      }
    sx += N / ncc; ex += N / ncc; sy += N / ncc; ey += N / ncc;
  }
  const float v = acc / (float)(ncc * px * px);
  if (cm) cm[ia] = v;
  if (sum) sum[ia] += (double)v;
}

// Closes one simulation of ccp's basic observables (simobs_b,
// sostat.cu:48-53,214-235): avg = sum / n_dp, m1 += avg, m2 += avg^2, and
// the sum restarts. FP64 with no fused operations, as ccp's host loop.
__global__ __launch_bounds__(BEAD_BLOCK) void k_close_sums(double *sum, double *m1, double *m2, uint32_t n,
                                                           double n_dp) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * BEAD_BLOCK + threadIdx.x;
  if (i >= n) return;
  const double avg = sum[i] / n_dp;
  m1[i] += avg;
  m2[i] += avg * avg;
  sum[i] = 0.0;
}

inline unsigned blocks_for(int n, int bs) { return (unsigned)((n + bs - 1) / bs); }

} // namespace

void launch_bin(hipStream_t s, const KParams &p, const GridGeom &g, int n_obj, const float4 *pos, bool het_only,
                uint32_t *count, uint32_t *blocksum, float4 *bucket, uint32_t *flags, uint32_t *hblocksum) {
  (void)p;
  if (n_obj <= 0) return;
  k_bin<<<blocks_for(n_obj, BEAD_BLOCK), BEAD_BLOCK, 0, s>>>(g, n_obj, pos, het_only ? 1 : 0, count, blocksum,
                                                            bucket, flags, hblocksum);
}

void launch_order(hipStream_t s, const OrderJob &a, const OrderJob *b, uint32_t *flags, uint64_t *step_inc,
                  const NoiseJob *noise, const KParams *bp, const BondJob *bond) {
  const OrderJob jb = b ? *b : a;
  const unsigned blocks = (unsigned)(a.g.nblk + (b ? b->g.nblk : 0));
  NoiseJob nj{};
  KParams p{};
  BondJob bj{};
  if (noise) nj = *noise;
  if (bond && bp) {
    p = *bp;
    bj = *bond;
  }
  k_order<<<blocks, SCAN_BLOCK, 0, s>>>(a, jb, flags, step_inc, nj, p, bj);
}

__global__ __launch_bounds__(64) void k_signal(const uint32_t *flags, uint32_t *host, uint32_t seq) {
  if (threadIdx.x != 0) return;
  for (int k = 0; k < FLAG_COUNT; ++k) host[1 + k] = flags[k];
  // the flags reach host memory before the sequence number does
  __threadfence_system();
  host[0] = seq;
}

void launch_signal(hipStream_t s, const uint32_t *flags, uint32_t *host, uint32_t seq) {
  k_signal<<<1, 64, 0, s>>>(flags, host, seq);
}

void launch_bead_terms(hipStream_t s, const NoiseJob *noise, const KParams &p, const BondJob *bond) {
  NoiseJob nj{};
  BondJob bj{};
  if (noise) nj = *noise;
  if (bond) bj = *bond;
  const uint32_t n = nj.z ? nj.n : bj.out ? bj.n : 0u; // both cover every bead
  if (n == 0) return;
  k_bead_terms<<<blocks_for((int)n, BEAD_BLOCK), BEAD_BLOCK, 0, s>>>(nj, p, bj);
}

template <int MODE, bool SPLIT, bool BATCH, int R, int LANES>
static void force_variant(hipStream_t s, int stage, const KParams &p, const KPtrs &q) {
  const unsigned g = blocks_for(p.N, Lanes<LANES>::BEADS_PER_BLOCK);
  if (stage == 0) k_force<MODE, 0, SPLIT, BATCH, R, LANES><<<g, BEAD_BLOCK, 0, s>>>(p, q);
  else if (stage == 1) k_force<MODE, 1, SPLIT, BATCH, R, LANES><<<g, BEAD_BLOCK, 0, s>>>(p, q);
  else k_force<MODE, 2, SPLIT, BATCH, R, LANES><<<g, BEAD_BLOCK, 0, s>>>(p, q);
}

template <int MODE, bool BATCH, int R>
static bool force_balanced(hipStream_t s, int stage, const KParams &p, const KPtrs &q, const NoiseJob *nz) {
  if constexpr (R == 2) {
    if (p.typed && q.work) {
      // het beads in het_blocks workgroups of 16, the others in workgroups of
      // 64, then (stage 2, when asked) the next step's normals, 128 beads per
      // workgroup
      NoiseJob z{};
      if (stage == 2 && nz && nz->z) z = *nz;
      const unsigned gt = (unsigned)p.het_blocks + blocks_for(p.N - p.n_het, T_BEADS_PER_BLOCK) +
                          (z.z ? blocks_for((int)z.n, BAL_BLOCK) : 0u);
      if (stage == 0) k_force_typed<MODE, 0, BATCH><<<gt, BAL_BLOCK, 0, s>>>(p, q, z);
      else if (stage == 1) k_force_typed<MODE, 1, BATCH><<<gt, BAL_BLOCK, 0, s>>>(p, q, z);
      else k_force_typed<MODE, 2, BATCH><<<gt, BAL_BLOCK, 0, s>>>(p, q, z);
      return z.z != nullptr;
    }
  }
  const unsigned g = blocks_for(p.N, C_BEADS_PER_BLOCK);
  if (q.fpair) {
    // pair sum with eight lanes per bead, then bonds, wall and the stage with one
    k_force_bal<MODE, 0, BATCH, R, false><<<g, BAL_BLOCK, 0, s>>>(p, q);
    const unsigned gi = blocks_for(p.N, BEAD_BLOCK);
    if (stage == 0) k_integrate<MODE, 0, BATCH><<<gi, BEAD_BLOCK, 0, s>>>(p, q);
    else if (stage == 1) k_integrate<MODE, 1, BATCH><<<gi, BEAD_BLOCK, 0, s>>>(p, q);
    else k_integrate<MODE, 2, BATCH><<<gi, BEAD_BLOCK, 0, s>>>(p, q);
    return false;
  }
  if (p.rounds == 4) {
    const unsigned g4 = blocks_for(p.N, C_BEADS_PER_BLOCK * 4);
    if (stage == 0) k_force_bal<MODE, 0, BATCH, R, true, 4><<<g4, BAL_BLOCK, 0, s>>>(p, q);
    else if (stage == 1) k_force_bal<MODE, 1, BATCH, R, true, 4><<<g4, BAL_BLOCK, 0, s>>>(p, q);
    else k_force_bal<MODE, 2, BATCH, R, true, 4><<<g4, BAL_BLOCK, 0, s>>>(p, q);
    return false;
  }
  if (p.rounds == 2) {
    const unsigned g2 = blocks_for(p.N, C_BEADS_PER_BLOCK * 2);
    if (stage == 0) k_force_bal<MODE, 0, BATCH, R, true, 2><<<g2, BAL_BLOCK, 0, s>>>(p, q);
    else if (stage == 1) k_force_bal<MODE, 1, BATCH, R, true, 2><<<g2, BAL_BLOCK, 0, s>>>(p, q);
    else k_force_bal<MODE, 2, BATCH, R, true, 2><<<g2, BAL_BLOCK, 0, s>>>(p, q);
    return false;
  }
  if (p.rounds == 8) {
    const unsigned g8 = blocks_for(p.N, C_BEADS_PER_BLOCK * 8);
    if (stage == 0) k_force_bal<MODE, 0, BATCH, R, true, 8><<<g8, BAL_BLOCK, 0, s>>>(p, q);
    else if (stage == 1) k_force_bal<MODE, 1, BATCH, R, true, 8><<<g8, BAL_BLOCK, 0, s>>>(p, q);
    else k_force_bal<MODE, 2, BATCH, R, true, 8><<<g8, BAL_BLOCK, 0, s>>>(p, q);
    return false;
  }
  if (stage == 0) k_force_bal<MODE, 0, BATCH, R><<<g, BAL_BLOCK, 0, s>>>(p, q);
  else if (stage == 1) k_force_bal<MODE, 1, BATCH, R><<<g, BAL_BLOCK, 0, s>>>(p, q);
  else k_force_bal<MODE, 2, BATCH, R><<<g, BAL_BLOCK, 0, s>>>(p, q);
  return false;
}

// kernel shape: cell reach 1 or 2; one lane per bead, one per plane, or
// eight sharing a compacted candidate list
template <int MODE, bool BATCH>
static bool force_shape(hipStream_t s, int stage, const KParams &p, const KPtrs &q, const NoiseJob *nz) {
  if (p.lanes == C_LANES) {
    if (p.reach == 2) return force_balanced<MODE, BATCH, 2>(s, stage, p, q, nz);
    return force_balanced<MODE, BATCH, 1>(s, stage, p, q, nz);
  }
  if (p.reach == 2) {
    if (p.lanes == 1) force_variant<MODE, false, BATCH, 2, 1>(s, stage, p, q);
    else force_variant<MODE, false, BATCH, 2, 5>(s, stage, p, q);
  } else {
    if (p.lanes == 1) force_variant<MODE, false, BATCH, 1, 1>(s, stage, p, q);
    else force_variant<MODE, false, BATCH, 1, 3>(s, stage, p, q);
  }
  return false;
}

bool launch_force(hipStream_t s, int mode, int stage, const KParams &p, const KPtrs &q, const NoiseJob *next_noise) {
  if (p.N <= 0) return false;
  if (q.sys) {
    // batches: the WCA/LJ set (or the relaxation of the initial conditions)
    // on one all-bead grid per system
    if (mode == MODE_ICG) return force_shape<MODE_ICG, true>(s, stage, p, q, next_noise);
    return force_shape<MODE_LJ, true>(s, stage, p, q, next_noise);
  }
  if (p.split && mode != MODE_ICG) {
    // short pairs on the fine grid, het-het tail on the het grid (reach 1, a lane per plane)
    if (mode == MODE_CCP) force_variant<MODE_CCP, true, false, 1, 3>(s, stage, p, q);
    else force_variant<MODE_LJ, true, false, 1, 3>(s, stage, p, q);
    return false;
  }
  if (mode == MODE_ICG) return force_shape<MODE_ICG, false>(s, stage, p, q, next_noise);
  if (mode == MODE_CCP) return force_shape<MODE_CCP, false>(s, stage, p, q, next_noise);
  return force_shape<MODE_LJ, false>(s, stage, p, q, next_noise);
}

void launch_contacts_batch(hipStream_t s, const SysParams *sys, const uint32_t *sys_slot, int n_sys, int n, float rc,
                           uint32_t bins, const float4 *spos, const uint32_t *cstart, uint32_t *maps) {
  if (n <= 0) return;
  k_contacts_batch<<<blocks_for(n, BEAD_BLOCK), BEAD_BLOCK, 0, s>>>(sys, sys_slot, n_sys, n, rc, bins, spos, cstart,
                                                                    maps);
}

void launch_radial_batch(hipStream_t s, const SysParams *sys, const uint32_t *sys_slot, int n_sys, int n, double a,
                         int n_b, const float4 *pos, double *out) {
  if (n <= 0) return;
  k_radial_batch<<<blocks_for(n, BEAD_BLOCK), BEAD_BLOCK, 0, s>>>(sys, sys_slot, n_sys, n, a, n_b, pos, out);
}

void launch_contacts(hipStream_t s, const KParams &p, const float4 *spos, const uint32_t *cstart, uint32_t *map,
                     uint32_t slot0, uint32_t id0) {
  if (p.N <= 0) return;
  const int rad = (int)ceilf(p.contact_rc / p.gf.csl);
  k_contacts<<<blocks_for(p.N, BEAD_BLOCK), BEAD_BLOCK, 0, s>>>(p, spos, cstart, map, rad, slot0, id0);
}

void launch_overlaps(hipStream_t s, const KParams &p, const float4 *spos, const uint32_t *cstart, uint32_t slot0,
                     float dmin2, unsigned long long *count) {
  if (p.N <= 0) return;
  k_overlaps<<<blocks_for(p.N, BEAD_BLOCK), BEAD_BLOCK, 0, s>>>(p, spos, cstart, slot0, dmin2, count);
}

void launch_noise(hipStream_t s, const KParams &p, uint64_t step, float4 *out) {
  if (p.N <= 0) return;
  k_noise<<<blocks_for(p.N, BEAD_BLOCK), BEAD_BLOCK, 0, s>>>(p, step, out);
}

void launch_observables(hipStream_t s, int N, double R_n, double nod, double a, int n_b, const float4 *pos,
                        double *out) {
  if (N <= 0) return;
  k_observables<<<blocks_for(N, BEAD_BLOCK), BEAD_BLOCK, (6 + 3 * n_b) * sizeof(double), s>>>(N, R_n, nod, a, n_b,
                                                                                               pos, out);
}

void launch_sd_cp(hipStream_t s, const float4 *pos, uint32_t N, uint32_t lsdcp, uint32_t i_s, uint32_t i_e,
                  float *sd, float *cp, double *sum_sd, double *sum_cp) {
  if (lsdcp == 0) return;
  k_sd_cp<<<blocks_for((int)lsdcp, BEAD_BLOCK), BEAD_BLOCK, 0, s>>>(pos, N, lsdcp, i_s, i_e, sd, cp, sum_sd, sum_cp);
}

void launch_ccp_cm(hipStream_t s, uint32_t N, uint32_t lcm, const float4 *pos, float *cm, double *sum) {
  if (lcm == 0) return;
  k_ccp_cm<<<blocks_for((int)lcm, BEAD_BLOCK), BEAD_BLOCK, 0, s>>>(N, lcm, pos, cm, sum);
}

void launch_bin_batch(hipStream_t s, const SysParams *sys, const uint32_t *sys_slot, int n_sys, int n,
                      const float4 *pos, uint32_t *count, uint32_t *blocksum, float4 *bucket, uint32_t *flags,
                      uint32_t *hblocksum) {
  if (n <= 0) return;
  k_bin_batch<<<blocks_for(n, BEAD_BLOCK), BEAD_BLOCK, 0, s>>>(sys, sys_slot, n_sys, n, pos, count, blocksum, bucket,
                                                              flags, hblocksum);
}

void launch_close_sums(hipStream_t s, double *sum, double *m1, double *m2, uint32_t n, double n_dp) {
  if (n == 0) return;
  k_close_sums<<<blocks_for((int)n, BEAD_BLOCK), BEAD_BLOCK, 0, s>>>(sum, m1, m2, n, n_dp);
}

} // namespace ccpmi
