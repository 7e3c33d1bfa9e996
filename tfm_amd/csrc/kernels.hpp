// Internal interface between the host side of libccpmi and its gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ccpmi {

// Bits of the 32-bit word stored in .w of every bead float4:
// bead index, type flags, and which of ccp's four bond vectors around the
// bead exist (calc_bf's lim_l/lim_u test, chrsim.cu:43-56).
constexpr uint32_t ID_BITS = 24; // the low 24 bits: v_mad_u32_u24 reads the index as it is
constexpr uint32_t ID_MASK = (1u << ID_BITS) - 1u;
constexpr uint32_t W_B3 = 1u << 26; // r[i+2] - r[i+1]
constexpr uint32_t W_LAD = 1u << 27;
constexpr uint32_t W_B0 = 1u << 28; // r[i-1] - r[i-2]
constexpr uint32_t W_B1 = 1u << 29; // r[i]   - r[i-1]
constexpr uint32_t W_B2 = 1u << 30; // r[i+1] - r[i]
constexpr uint32_t W_HET = 1u << 31; // the sign bit: a het partner is a negative word

// cells per scan block of the order kernel; each block reads the totals of
// the blocks before it, so fewer, larger blocks keep that prefix cheap
constexpr int SCAN_BLOCK = 256;
// threads per block of the bead kernels
constexpr int BEAD_BLOCK = 256;

// device error flags
enum : int { FLAG_OVERFLOW = 0, FLAG_OUTSIDE = 1, FLAG_NONFINITE = 2, FLAG_COUNT = 4 };

enum Mode : int { MODE_CCP = 0, MODE_ICG = 1, MODE_LJ = 2 };

// Geometry of one uniform cell grid centred on the nucleus (ccp's sugrid,
// sugrid.cuh:15-54): cell (gx, gy, gz) = floor(r / csl) + cps/2, index
// coff + gx + cps (gy + cps gz). coff places the grid of one system inside
// the concatenated cell range of a batch (0 for a single system).
struct GridGeom {
  float csl;
  float icsl; // 1/csl for the cell lookup of the step grids; 0: divide by csl (ccp's rule, export grid)
  int cps, half, n_c, cap, nblk, coff;
};

// Per-system parameters of a batch of independent systems that share one
// bead array and one concatenated cell range. Beads of system s hold the
// slots [slot, slot + n) of the cell-ordered arrays and the bead indices
// [slot, slot + n) of the index-ordered ones.
struct SysParams {
  float R_n, R_o, R_b, noc, boc, nod, bod;
  float eps_het, eps_lam;
  uint32_t seed_lo, seed_hi;
  uint32_t slot, n;
  GridGeom g;
  int frozen; // initial-condition relaxation done: zero force, the beads keep their positions
};

// Device buffers of one grid. count/blocksum/bucket receive the binning of
// the next positions; cstart/spos/soi hold the ordered cell list.
struct GridBufs {
  uint32_t *count;
  uint32_t *bs[2];   // per-scan-block totals, double buffered
  float4 *bucket;    // n_c * cap unsorted members (position, index in the w bits)
  uint32_t *cstart;  // n_c + 1 starts into spos
  float4 *spos;      // positions in cell order
  uint32_t *soi;     // bead index in cell order (optional)
};

struct KParams {
  int N, n_l;
  GridGeom gf; // grid of every bead (all pairs, or the short-range pass when split), LBS sites
  GridGeom gh; // grid of het beads (long-range het-het pass when split)
  int split;   // 1: short-range pairs on gf + het-het tail on gh; 0: every pair on gf
  int reach;   // neighbour cells scanned on each side: 1 (27 cells) or 2 (125 cells of half the side)
  int lanes;   // lanes per bead of the force kernel: 2 reach + 1 (one per z-plane), 1, or 8 (balanced list)
  int group;   // 8 lanes: within a workgroup, beads with the long (het) cutoff take the first wavefronts
  int rounds;  // 8 lanes: beads per lane group of a wavefront (1, 2, 4 or 8), eight per round of the pair phase
  // 8 lanes at reach 2 over a type-ordered work list (k_force_typed): the
  // het beads (n_het of them) in the first het_blocks workgroups, eight lanes
  // each; the others two lanes each over reach 1 of the same cells
  int typed, n_het, het_blocks;
  int frozen;  // batches: the bead's system is frozen (set per bead from its SysParams)
  // nucleus geometry (cngeom, chrdat.cuh:48-69)
  float R_n, R_o, R_b, noc, boc, nod, bod;
  // bonds and bending
  float k_e, l_0, k_b;
  int bend_all, bond_type;
  float fene_k, fene_r02;
  // Wang-Frenkel pair set (ccp)
  float e_p, rco, rco2, aco, aco2;
  // LBS binding: lfac_d = e_l 8/(3 lco^2) in FP64 as ccp evaluates it
  double lfac_d;
  float lco;
  // WCA/LJ set in units of sigma (sigma = 1)
  float eps_wca, rm2, eps_het, rc_lj2, eps_wall, eps_lam, rc_lam2;
  // squared candidate-list radii of the balanced kernel: cutoff + 1e-3 sigma
  // for het beads (the het-het cutoff) and for the others (the short one), in
  // the simulation's pair set (the relaxation set uses rco for every bead)
  float trim2_long, trim2_short;
  // integrator
  float dt, half_dt, sd;
  float fcap2; // squared force cap of the integrator (0 = off)
  uint32_t seed_lo, seed_hi;
  // contacts
  float contact_rc;
  int contact_bin_size, contact_bins;
};

struct KPtrs {
  const float4 *fspos;     // fine grid of the evaluated positions
  const uint32_t *fcstart;
  const float4 *hspos;     // het grid of the evaluated positions
  const uint32_t *hcstart;
  const float4 *lspos;     // LBS sites on the fine geometry
  const uint32_t *lstart;
  const float4 *pos;       // bead positions of the evaluated stage, by bead index
  float4 *out0;            // stage 1: er, stage 2: new r, stage 0: force
  float4 *pstar;           // r + f dt/2 + noise (written by stage 1, read by stage 2)
  uint32_t *fcount, *fblocksum; // binning of the output into the next fine grid
  uint32_t *fhblocksum;    // ... its per-block het totals (type-ordered work list; nullptr: none)
  float4 *fbucket;
  uint32_t *hcount, *hblocksum; // ... and into the next het grid
  float4 *hbucket;
  uint32_t *flags;         // FLAG_COUNT error counters
  const uint64_t *step;    // global step counter (noise counter)
  float4 *fpair;           // pair forces by bead index (balanced kernel with a separate integration pass)
  const float4 *noise;     // unit normal triples of this step by bead index (first Heun stage)
  const float4 *fbond;     // bonds + bending + wall per bead by bead index (balanced kernel; nullptr: inline)
  const uint32_t *work;    // type-ordered work list of the fine grid (k_force_typed)
  // batches only (sys == nullptr for a single system)
  const SysParams *sys;    // n_sys systems
  const uint32_t *sys_slot; // n_sys + 1 first slots
  int n_sys;
};

// One grid's ordering work inside a k_order launch.
struct OrderJob {
  GridGeom g;
  uint32_t *count;
  const uint32_t *bs_read;
  uint32_t *bs_clear;
  const float4 *bucket; // members' positions, index in the w bits
  float4 *spos;
  uint32_t *soi;
  uint32_t *cstart;
  // type-ordered work list (nullptr: none): the slots of het members first,
  // then the others, each part in slot order; per-block het totals of the
  // binning (read, and the other buffer cleared); n_het het beads of n_slots
  uint32_t *work = nullptr;
  const uint32_t *hbs_read = nullptr;
  uint32_t *hbs_clear = nullptr;
  uint32_t n_het = 0, n_slots = 0;
};

// Unit normal triples of one step for every bead, by bead index: Philox
// counter {local bead, step, attempt} under the bead's system key. The order
// pass after the first Heun stage fills them for the next step (ahead = 1).
struct NoiseJob {
  float4 *z;               // out: (z0, z1, z2, attempts) per bead
  const uint64_t *step;    // device step counter
  uint32_t ahead;          // noise of step *step + ahead
  uint32_t n;              // beads
  uint32_t seed_lo, seed_hi;
  const SysParams *sys;    // batch: per-system keys and first beads (nullptr: one system)
  const uint32_t *sys_slot;
  int n_sys;
};

// Bonded (bonds, bending) and wall forces of the positions the next force
// pass evaluates, one lane per bead, written by bead index. The order pass
// after each stage fills them for the next stage.
struct BondJob {
  float4 *out;              // out: force per bead (nullptr: none)
  const float4 *pos;        // positions by bead index
  uint32_t n;               // beads
  int mode;                 // pair set of the wall (MODE_*)
  const SysParams *sys;     // batch: per-system geometry (nullptr: one system)
  const uint32_t *sys_slot;
  int n_sys;
};

// Het beads per 128-lane workgroup of k_force_typed (eight lanes each).
constexpr int TYPED_HET_BEADS_PER_BLOCK = 16;

// Het count of the type-ordered work list and the workgroups of its het part
// (TYPED_HET_BEADS_PER_BLOCK beads each, rounded to a multiple of the 8 XCDs
// so that the other part starts on XCD 0).
inline void set_het_count(KParams &k, uint32_t n_het) {
  constexpr uint32_t per = (uint32_t)TYPED_HET_BEADS_PER_BLOCK;
  k.n_het = (int)n_het;
  k.het_blocks = (int)(((n_het + per - 1u) / per + 7u) / 8u * 8u);
}

// Launchers (kernels.hip). All are asynchronous on `s`.
void launch_bin(hipStream_t s, const KParams &p, const GridGeom &g, int n_obj, const float4 *pos, bool het_only,
                uint32_t *count, uint32_t *blocksum, float4 *bucket, uint32_t *flags, uint32_t *hblocksum = nullptr);
// The order pass also fills the next force pass's per-bead terms when given:
// normals (noise) and bonded + wall forces (bp, bond).
void launch_order(hipStream_t s, const OrderJob &a, const OrderJob *b, uint32_t *flags, uint64_t *step_inc,
                  const NoiseJob *noise = nullptr, const KParams *bp = nullptr, const BondJob *bond = nullptr);
// the same per-bead terms on their own (before the first stage of a run)
void launch_bead_terms(hipStream_t s, const NoiseJob *noise, const KParams &p, const BondJob *bond);
// Copy the FLAG_COUNT error flags into host[1..] and then seq into host[0]
// (host: device address of coherent pinned host memory), behind every
// launch ahead of it in the stream.
void launch_signal(hipStream_t s, const uint32_t *flags, uint32_t *host, uint32_t seq);
// next_noise (optional, stage 2): normals for the type-ordered kernel to draw
// in workgroups after its own, which run in the launch's tail; returns whether
// it took them (false: the caller draws them elsewhere).
bool launch_force(hipStream_t s, int mode, int stage, const KParams &p, const KPtrs &q,
                  const NoiseJob *next_noise = nullptr);
// Contacts of the beads in slots [slot0, slot0 + p.N) (bead indices from
// id0) on p.gf, binned by (index - id0) / contact_bin_size.
void launch_contacts(hipStream_t s, const KParams &p, const float4 *spos, const uint32_t *cstart, uint32_t *map,
                     uint32_t slot0 = 0, uint32_t id0 = 0);
// slots [slot0, slot0 + p.N) of spos, pairs counted on p.gf
void launch_overlaps(hipStream_t s, const KParams &p, const float4 *spos, const uint32_t *cstart, uint32_t slot0,
                     float dmin2, unsigned long long *count);
void launch_noise(hipStream_t s, const KParams &p, uint64_t step, float4 *out);
void launch_observables(hipStream_t s, int N, double R_n, double nod, double a, int n_b, const float4 *pos,
                        double *out);
void launch_sd_cp(hipStream_t s, const float4 *pos, uint32_t N, uint32_t lsdcp, uint32_t i_s, uint32_t i_e,
                  float *sd, float *cp, double *sum_sd, double *sum_cp);
void launch_ccp_cm(hipStream_t s, uint32_t N, uint32_t lcm, const float4 *pos, float *cm, double *sum);
void launch_bin_batch(hipStream_t s, const SysParams *sys, const uint32_t *sys_slot, int n_sys, int n,
                      const float4 *pos, uint32_t *count, uint32_t *blocksum, float4 *bucket, uint32_t *flags,
                      uint32_t *hblocksum = nullptr);
// every system of a batch in one launch: contact maps (maps + s bins^2) and
// the all-bead radial density (out + s n_b, FP64 sums)
void launch_contacts_batch(hipStream_t s, const SysParams *sys, const uint32_t *sys_slot, int n_sys, int n, float rc,
                           uint32_t bins, const float4 *spos, const uint32_t *cstart, uint32_t *maps);
void launch_radial_batch(hipStream_t s, const SysParams *sys, const uint32_t *sys_slot, int n_sys, int n, double a,
                         int n_b, const float4 *pos, double *out);
void launch_close_sums(hipStream_t s, double *sum, double *m1, double *m2, uint32_t n, double n_dp);

} // namespace ccpmi
