// Host side of libccpmi: the C ABI of include/ccpmi.h.
//
// Parameters, validation messages and file formats follow ccp
// (Program/src/chrdat.cu, chrsim.cu, util.cu); the step runs on the gfx950
// kernels of kernels.hip, captured into hipGraphs of `graph_steps` steps.

#include "ccpmi.h"
#include "sim.hpp"
#include "stats.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

using namespace ccpmi;
using namespace ccpmi::host;


void ccp_sim::configure(const ParMap &par) {
  // chrdat::chrdat and cngeom (chrdat.cu:13-75), chrsim::chrsim (chrsim.cu:421-439)
  N = par.get<uint32_t>("number_of_particles", N_DEF);
  R_n = par.get<float>("nucleus_radius", 0.0f);
  R_o = par.get<float>("opening_radius", 0.0f);
  R_b = par.get<float>("bleb_radius", 0.0f);
  noc = (float)std::cos(std::asin((double)R_o / R_n));
  boc = (float)std::cos(std::asin((double)R_o / R_b));
  nod = (float)std::sqrt((double)R_n * R_n - (double)R_o * R_o);
  bod = (float)std::sqrt((double)R_b * R_b - (double)R_o * R_o);
  d_m = (nod + bod) + R_b;
  if (!(0.0f < R_n && R_n < 1000.0f)) throw Error(CCP_ERR_PARAM, "nucleus_radius out of range");
  if (!(0.0f <= R_o && R_o < R_n)) throw Error(CCP_ERR_PARAM, "opening_radius out of range");
  if (!(R_o <= R_b && R_b < R_n)) throw Error(CCP_ERR_PARAM, "bleb_radius out of range");
  record("R_n = " + cnfs(R_n, 5, '0', 2) + " R_o = " + cnfs(R_o, 5, '0', 2) + " R_b = " + cnfs(R_b, 5, '0', 2) +
         " ");

  T = par.get<float>("temperature", 298.0f);
  n_l = par.get<uint32_t>("number_of_lbs", 0u);
  fpf = par.get<uint32_t>("frames_per_file", 128u);
  if (!(1 <= N && N < (1u << ID_BITS))) throw Error(CCP_ERR_PARAM, "number_of_particles out of range");
  if (!(0.0f <= T && T < 1000.0f)) throw Error(CCP_ERR_PARAM, "temperature out of range");
  if (!(n_l < 100000)) throw Error(CCP_ERR_PARAM, "number_of_lbs out of range");
  if (!(1 <= fpf && fpf < 10000)) throw Error(CCP_ERR_PARAM, "frames_per_file out of range");
  float cvf = (float)(N * std::pow((RCO / 2.0) / (R_n - (RCO / 2.0)), 3.0));
  if (cvf > 0.4f) throw Error(CCP_ERR_PARAM, "chromatin volume fraction above 0.4");
  float noacf = (float)(2.0 / (1.0 + noc));
  float laf = (float)(noacf * n_l * std::pow(LCO / (R_n - RCO), 2.0) / 4.0);
  if (laf > 0.4f) throw Error(CCP_ERR_PARAM, "lbs area fraction above 0.4");
  record("N = " + cnfs(N, 5, '0') + " T = " + cnfs(T, 5, '0', 1) + " n_l = " + cnfs(n_l, 5, '0') +
         " fpf = " + cnfs(fpf, 4, '0') + " ");
  record("cvf = " + cnfs(cvf, 5, '0', 3) + " laf = " + cnfs(laf, 5, '0', 3) + " ");

  dt = par.get<float>("timestep", DT_DEF);
  if (!(dt > 0.0f && dt < 1.0f)) throw Error(CCP_ERR_PARAM, "timestep out of range");
  spf = par.get<uint32_t>("steps_per_frame", (uint32_t)(1.0 / dt));
  tpb = par.get<uint32_t>("threads_per_block", 128u);
  if (!(1 <= spf && spf < 10000)) throw Error(CCP_ERR_PARAM, "steps_per_frame out of range");
  if (!(1 <= tpb && tpb < 1025)) throw Error(CCP_ERR_PARAM, "threads_per_block out of range");
  record("spf = " + cnfs(spf, 4, '0') + " tpb = " + cnfs(tpb, 4, '0') + " ");
  sd = (float)std::sqrt(2.0 * K_B * (double)T * (double)dt);

  // ---- build-defined keys
  seed = par.has("seed") ? par.get<uint64_t>("seed", 0) : (uint64_t)time(nullptr);
  replica = par.get<uint32_t>("replica", 0u);
  if (replica >= (1u << 15)) throw Error(CCP_ERR_PARAM, "replica out of range");
  std::string ffs = par.get<std::string>("force_field", "ccp");
  if (ffs == "ccp") ff = CCP_FF_CCP;
  else if (ffs == "lj") ff = CCP_FF_LJ;
  else throw Error(CCP_ERR_PARAM, "force_field must be ccp or lj");
  const bool lj = ff == CCP_FF_LJ;
  const float rc_lj = 2.5f;
  eps_het = par.get<float>("epsilon_het", 0.5f);
  eps_lam = par.get<float>("epsilon_lamina", 1.0f);
  k_b = par.get<float>("bending_constant", lj ? 2.0f : K_BEND);
  std::string bend = par.get<std::string>("bending", lj ? "all" : "het");
  if (bend != "all" && bend != "het") throw Error(CCP_ERR_PARAM, "bending must be het or all");
  bend_all = bend == "all";
  std::string bond = par.get<std::string>("bond", "harmonic");
  if (bond != "harmonic" && bond != "fene") throw Error(CCP_ERR_PARAM, "bond must be harmonic or fene");
  bond_type = bond == "fene";
  contact_rc = par.get<float>("contact_cutoff", lj ? 1.5f : ACO);
  float cutoff = lj ? rc_lj : ACO;
  cutoff = std::max(std::max(cutoff, LCO), contact_rc);
  // cell_reach 2 scans 5 x 5 x 5 cells of at least half the cutoff instead
  // of ccp's 3 x 3 x 3 (the same pairs, fewer distant candidates)
  reach = par.get<uint32_t>("cell_reach", 1u);
  if (reach != 1 && reach != 2) throw Error(CCP_ERR_PARAM, "cell_reach out of range");
  csl = par.get<float>("cell_size", (lj ? rc_lj : ACO) / reach);
  if (!(csl * reach >= cutoff)) throw Error(CCP_ERR_PARAM, "cell_size below the interaction cutoff");
  // fine grid: side >= every short-range cutoff (pair repulsion, LBS
  // binding, the initial-overlap distance)
  // (the split is off when fine_cell_size equals cell_size, the default)
  const float short_cut = lj ? (float)std::pow(2.0, 1.0 / 6.0) : RCO;
  csl_f = par.get<float>("fine_cell_size", csl);
  if (csl_f > csl) csl_f = csl;
  split = csl_f < csl;
  if (split && !(csl_f >= short_cut && csl_f >= LCO && csl_f >= MIS && csl_f >= contact_rc / 2.0f))
    throw Error(CCP_ERR_PARAM, "fine_cell_size below the short-range cutoff");
  // the overlap count of the initial condition scans one cell around each bead
  if (!(csl_f >= MIS)) throw Error(CCP_ERR_PARAM, "cell_size below the overlap distance");
  if (split && reach != 1) throw Error(CCP_ERR_PARAM, "fine_cell_size needs cell_reach 1");
  uint32_t cap = par.get<uint32_t>("cell_capacity", 32u);
  uint32_t cap_f = par.get<uint32_t>("fine_cell_capacity", 16u);
  if (!(1 <= cap && cap <= 1024)) throw Error(CCP_ERR_PARAM, "cell_capacity out of range");
  if (!(1 <= cap_f && cap_f <= 1024)) throw Error(CCP_ERR_PARAM, "fine_cell_capacity out of range");
  auto geometry = [&](float side, uint32_t capacity) {
    GridGeom g{};
    uint32_t cps = 2 * (uint32_t)std::ceil(d_m / side);
    if (cps < 4) cps = 4;
    if ((uint64_t)cps * cps * cps >= (1ull << 31)) throw Error(CCP_ERR_PARAM, "cell_size too small for the nucleus");
    g.csl = side;
    g.icsl = 1.0f / side;
    g.cps = (int)cps;
    g.half = (int)(cps / 2);
    g.n_c = (int)(cps * cps * cps);
    g.cap = (int)capacity;
    g.nblk = (g.n_c + SCAN_BLOCK - 1) / SCAN_BLOCK;
    return g;
  };
  gh.g = geometry(csl, cap);
  gx.g = gh.g;
  gx.g.icsl = 0.0f; // the exported cell list divides by csl as ccp does
  gf.g = geometry(csl_f, split ? cap_f : cap);
  lg = gf.g;
  fcap = par.get<float>("force_cap", lj ? 0.2f / dt : 0.0f);
  if (!(fcap >= 0.0f)) throw Error(CCP_ERR_PARAM, "force_cap out of range");
  lanes = par.get<uint32_t>("lanes_per_bead", 2 * reach + 1);
  if (lanes != 1 && lanes != 2 * reach + 1 && lanes != 8) throw Error(CCP_ERR_PARAM, "lanes_per_bead out of range");
  // the balanced kernel scans one grid: the split pair sum has no 8-lane form
  if (lanes == 8 && split) throw Error(CCP_ERR_PARAM, "lanes_per_bead 8 excludes fine_cell_size");
  // with 8 lanes per bead: bonds, wall and the stage in a second, one-lane-per-bead
  // pass (off by default: measured 15 % slower at N = 131072)
  integrate_pass = par.get<uint32_t>("integrate_pass", 0u) != 0;
  // with 8 lanes per bead: reorder each workgroup's beads by cutoff so that a
  // wavefront's lists have similar lengths (no change to any result)
  group_types = par.get<uint32_t>("group_types", 1u) != 0;
  // with 8 lanes per bead: a wavefront runs the pair phase for `bal_rounds`
  // groups of eight beads, then the per-bead work of all of them at once
  rounds = par.get<uint32_t>("bal_rounds", 1u);
  if (rounds != 1 && rounds != 2 && rounds != 4 && rounds != 8) throw Error(CCP_ERR_PARAM, "bal_rounds out of range");
  // with 8 lanes at reach 2: the force kernel takes the beads from a
  // type-ordered work list (het beads first, eight lanes each; the others
  // two lanes each over reach 1), when the short cutoffs of the pair set and
  // of the relaxation stay within one cell side
  {
    const double margin = 1e-3, rs = lj ? std::pow(2.0, 1.0 / 6.0) : RCO;
    const double reach1 = std::max(rs, (double)RCO) + margin;
    typed = par.get<uint32_t>("type_order", 1u) != 0 && lanes == 8 && reach == 2 && !split && !integrate_pass &&
            rounds == 1 && reach1 < csl_f;
  }
  if (rounds > 1 && integrate_pass) throw Error(CCP_ERR_PARAM, "bal_rounds > 1 excludes integrate_pass");
  graph_steps = par.get<uint32_t>("graph_steps", 64u);
  if (!(1 <= graph_steps && graph_steps <= 4096)) throw Error(CCP_ERR_PARAM, "graph_steps out of range");
  use_graphs = par.get<uint32_t>("graphs", 1u) != 0;

  // chains: ccp's chromosome limits when N = 30386 (chrdat.cuh:27-34),
  // otherwise number_of_chains equal parts or an explicit chain_lengths list
  chain_off.clear();
  if (par.has("chain_lengths")) {
    std::string s = par.get<std::string>("chain_lengths", "");
    std::replace(s.begin(), s.end(), ',', ' ');
    std::istringstream is(s);
    chain_off.push_back(0);
    long long L;
    while (is >> L) {
      if (L <= 0) throw Error(CCP_ERR_PARAM, "chain_lengths out of range");
      chain_off.push_back(chain_off.back() + (uint32_t)L);
    }
    if (chain_off.back() != N || chain_off.size() < 2) throw Error(CCP_ERR_PARAM, "chain_lengths must sum to N");
  } else {
    uint32_t nch = par.get<uint32_t>("number_of_chains", 0u);
    if (nch == 0 && N == N_DEF) {
      chain_off.assign(CHR_LIM, CHR_LIM + 13);
    } else {
      if (nch == 0) nch = 1;
      if (nch > N) throw Error(CCP_ERR_PARAM, "number_of_chains out of range");
      chain_off.push_back(0);
      for (uint32_t c = 0; c < nch; ++c) chain_off.push_back(chain_off.back() + N / nch + (c < N % nch ? 1 : 0));
    }
  }

  seq_file = par.get<std::string>("sequence_file", "Program/data/sequence.txt");
  std::string seq = par.get<std::string>("sequence", lj ? "het_lad" : "ccp");
  if (seq != "ccp" && seq != "het_lad") throw Error(CCP_ERR_PARAM, "sequence must be ccp or het_lad");
  seq_mode = seq == "het_lad";
  het_fraction = par.get<float>("het_fraction", 0.3f);
  het_mean = par.get<float>("het_block_mean", 20.0f);
  lad_fraction = par.get<float>("lad_fraction", 0.1f);
  std::string lp = par.get<std::string>("lad_placement", "ends");
  if (lp != "ends" && lp != "uniform") throw Error(CCP_ERR_PARAM, "lad_placement must be ends or uniform");
  lad_ends = lp == "ends";
  if (!(het_fraction > 0.0f && het_fraction < 1.0f)) throw Error(CCP_ERR_PARAM, "het_fraction out of range");
  if (!(lad_fraction >= 0.0f && lad_fraction < 1.0f)) throw Error(CCP_ERR_PARAM, "lad_fraction out of range");
  if (!(het_mean >= 1.0f)) throw Error(CCP_ERR_PARAM, "het_block_mean out of range");

  cbin_size = par.get<uint32_t>("contact_bin_size", std::max<uint32_t>(1u, (N + 1023) / 1024));
  if (cbin_size < 1) throw Error(CCP_ERR_PARAM, "contact_bin_size out of range");
  cbins = (N + cbin_size - 1) / cbin_size;
  if ((uint64_t)cbins * cbins > (1ull << 28)) throw Error(CCP_ERR_PARAM, "contact map too large");
  cinterval = par.get<uint32_t>("contact_interval", 0u);

  hr.assign(3 * (size_t)N, 0.0f);
  hlr.assign(3 * (size_t)n_l, 0.0f);
  hfl.assign(N, 0);
}

void ccp_sim::fill_kparams() {
  KParams &k = kp;
  k.N = (int)N;
  k.n_l = (int)n_l;
  k.gf = gf.g;
  k.gh = gh.g;
  if (!split) k.gh.n_c = 0;
  k.split = split ? 1 : 0;
  k.lanes = (int)lanes;
  k.group = group_types ? 1 : 0;
  k.rounds = (int)rounds;
  k.typed = typed ? 1 : 0;
  set_het_count(k, 0);
  k.reach = split ? 1 : (int)reach;
  k.R_n = R_n; k.R_o = R_o; k.R_b = R_b;
  k.noc = noc; k.boc = boc; k.nod = nod; k.bod = bod;
  const bool lj = ff == CCP_FF_LJ;
  k.k_e = K_E;
  k.l_0 = L_0;
  k.k_b = k_b;
  k.bend_all = bend_all;
  k.bond_type = bond_type;
  k.fene_k = 30.0f;
  k.fene_r02 = 1.5f * 1.5f;
  k.e_p = E_P;
  k.rco = RCO;
  k.rco2 = RCO * RCO;
  k.aco = ACO;
  k.aco2 = ACO * ACO;
  // ccp's e_l * (8.0 / (3.0 * lco * lco)): float constants, double literals
  k.lfac_d = (double)E_L * (8.0 / (3.0 * (double)LCO * (double)LCO));
  k.lco = LCO;
  k.eps_wca = 1.0f;
  k.rm2 = (float)std::pow(2.0, 1.0 / 3.0);
  k.eps_het = eps_het;
  k.rc_lj2 = 2.5f * 2.5f;
  k.eps_wall = 1.0f;
  k.eps_lam = eps_lam;
  k.rc_lam2 = 2.5f * 2.5f;
  {
    const double margin = 1e-3; // covers the rounding of cell edges
    const double rl = lj ? 2.5 : ACO, rs = lj ? std::pow(2.0, 1.0 / 6.0) : RCO;
    k.trim2_long = (float)((rl + margin) * (rl + margin));
    k.trim2_short = (float)((rs + margin) * (rs + margin));
  }
  k.dt = dt;
  k.half_dt = 0.5f * dt;
  k.sd = sd;
  k.fcap2 = fcap * fcap;
  // replicas > 0 draw their noise with another key (golden-ratio offset)
  const uint64_t key = seed + (uint64_t)replica * 0x9E3779B97F4A7C15ull;
  k.seed_lo = (uint32_t)key;
  k.seed_hi = (uint32_t)(key >> 32);
  k.contact_rc = contact_rc;
  k.contact_bin_size = (int)cbin_size;
  k.contact_bins = (int)cbins;
}

void ccp_sim::allocate() {
  hip_check(hipGetDevice(&dev));
  hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  hip_check(hipHostMalloc(reinterpret_cast<void **>(&hsig), 64, hipHostMallocMapped | hipHostMallocCoherent));
  std::memset(hsig, 0, 64);
  hip_check(hipHostGetDevicePointer(reinterpret_cast<void **>(&hsig_dev), hsig, 0));
  r.alloc(N); er.alloc(N); pstar.alloc(N); aux.alloc(N); znoise.alloc(N);
  if (bond_pass()) fbond.alloc(N);
  if (lanes == 8 && integrate_pass) fpair.alloc(N);
  lr.alloc(n_l); lspos.alloc(n_l); lstart.alloc(gf.g.n_c + 1);
  flags.alloc(FLAG_COUNT);
  cmap.alloc((size_t)cbins * cbins);
  dstep.alloc(1);
  dcount.alloc(1);
  gf.alloc(std::max(N, n_l), typed);
  if (split) gh.alloc(N);
  hip_check(hipMemset(flags.p, 0, FLAG_COUNT * sizeof(uint32_t)));
  hip_check(hipMemset(cmap.p, 0, (size_t)cbins * cbins * sizeof(uint32_t)));
  hip_check(hipMemset(dstep.p, 0, sizeof(uint64_t)));
  hip_check(hipMemset(lstart.p, 0, (gf.g.n_c + 1) * sizeof(uint32_t)));
}

uint32_t ccp_sim::wbits(uint32_t i) const {
  uint32_t w = i;
  if (hfl[i] & CCP_HET) w |= W_HET;
  if (hfl[i] & CCP_LAD) w |= W_LAD;
  auto it = std::upper_bound(chain_off.begin(), chain_off.end(), i);
  uint32_t s = *(it - 1), e = *it;
  if (i >= s + 2) w |= W_B0;
  if (i >= s + 1) w |= W_B1;
  if (i + 1 <= e - 1) w |= W_B2;
  if (i + 2 <= e - 1) w |= W_B3;
  return w;
}

void ccp_sim::upload_positions() {
  uint32_t n_het = 0;
  for (uint32_t i = 0; i < N; ++i) n_het += (hfl[i] & CCP_HET) ? 1u : 0u;
  if (n_het != (uint32_t)kp.n_het) {
    // captured graphs hold the kernel parameters by value
    drop_graphs();
    set_het_count(kp, n_het);
  }
  std::vector<float4> h(N);
  for (uint32_t i = 0; i < N; ++i) {
    uint32_t w = wbits(i);
    float wf;
    std::memcpy(&wf, &w, 4);
    h[i] = make_float4(hr[3 * i], hr[3 * i + 1], hr[3 * i + 2], wf);
  }
  hip_check(hipMemcpyAsync(r.p, h.data(), N * sizeof(float4), hipMemcpyHostToDevice, stream));
  hip_check(hipStreamSynchronize(stream));
  invalidate();
}

void ccp_sim::upload_lbs() {
  std::vector<float4> h(std::max<uint32_t>(n_l, 1));
  for (uint32_t l = 0; l < n_l; ++l) {
    float wf;
    std::memcpy(&wf, &l, 4);
    h[l] = make_float4(hlr[3 * l], hlr[3 * l + 1], hlr[3 * l + 2], wf);
  }
  if (n_l) hip_check(hipMemcpyAsync(lr.p, h.data(), n_l * sizeof(float4), hipMemcpyHostToDevice, stream));
  hip_check(hipStreamSynchronize(stream));
  lgrid_valid = false;
}

void ccp_sim::download_positions() {
  std::vector<float4> h(N);
  hip_check(hipMemcpyAsync(h.data(), r.p, N * sizeof(float4), hipMemcpyDeviceToHost, stream));
  hip_check(hipStreamSynchronize(stream));
  for (uint32_t i = 0; i < N; ++i) {
    hr[3 * i] = h[i].x;
    hr[3 * i + 1] = h[i].y;
    hr[3 * i + 2] = h[i].z;
  }
}

// Waits for the stream's work and checks the error flags: k_signal, queued
// behind that work, writes the flags and then a new sequence number into
// pinned host memory, and the host polls for the number (a few microseconds
// fewer than a copy into host memory followed by an event wait). Every few
// thousand polls the stream is queried, so that a failed launch is reported
// instead of waited for.
void ccp_sim::check_flags() {
  const uint32_t seq = ++sig_seq;
  launch_signal(stream, flags.p, hsig_dev, seq);
  hip_check(hipGetLastError());
  volatile uint32_t *sig = hsig;
  for (uint32_t k = 1; sig[0] != seq; ++k) {
    if ((k & 4095u) == 0u) {
      const hipError_t e = hipStreamQuery(stream);
      if (e == hipSuccess) {
        // the stream is done: the signal's host writes may still be on their
        // way, so wait a bounded while longer before calling it lost
        for (uint32_t w = 0; sig[0] != seq; ++w)
          if (w > (1u << 26)) throw Error(CCP_ERR_HIP, "stream drained without its completion signal");
        break;
      }
      if (e != hipErrorNotReady) hip_check(e);
    }
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  uint32_t f[FLAG_COUNT];
  for (int k = 0; k < FLAG_COUNT; ++k) f[k] = sig[1 + k];
  if (f[FLAG_OVERFLOW] || f[FLAG_OUTSIDE] || f[FLAG_NONFINITE]) {
    hip_check(hipMemset(flags.p, 0, sizeof(f)));
    if (f[FLAG_NONFINITE]) throw Error(CCP_ERR_STATE, "non-finite bead position");
    if (f[FLAG_OUTSIDE]) throw Error(CCP_ERR_STATE, "bead left the cell grid");
    throw Error(CCP_ERR_STATE, "cell_capacity exceeded; raise cell_capacity");
  }
}

// Bin n_obj positions with G's buffers in geometry `geom` and order them
// into (out_spos, out_cstart[, out_soi]).
void ccp_sim::build_grid(HostGrid &G, const GridGeom &geom, const float4 *pos, int n_obj, bool het_only,
                         float4 *out_spos, uint32_t *out_cstart, uint32_t *out_soi, bool work_list) {
  const bool tw = work_list && G.work.p;
  launch_bin(stream, kp, geom, n_obj, pos, het_only, G.count.p, G.bs(G.clean), G.bucket.p, flags.p,
             tw ? G.hbs(G.clean) : nullptr);
  const OrderJob j = G.order_job(geom, tw, (uint32_t)kp.n_het, (uint32_t)n_obj, out_spos, out_cstart, out_soi);
  launch_order(stream, j, nullptr, flags.p, nullptr);
  hip_check(hipGetLastError());
  G.clean ^= 1;
}

// Overflow flag of the last builds (synchronizes the stream); clears it.
bool ccp_sim::take_overflow() {
  uint32_t f = 0;
  hip_check(hipMemcpyAsync(&f, flags.p + FLAG_OVERFLOW, 4, hipMemcpyDeviceToHost, stream));
  hip_check(hipStreamSynchronize(stream));
  if (f) hip_check(hipMemset(flags.p + FLAG_OVERFLOW, 0, 4));
  return f != 0;
}

// Double every grid's cell capacity (bucket storage); captured graphs refer
// to the old buffers and are dropped.
void ccp_sim::grow_capacity() {
  if (gh.g.cap >= 4096 || gf.g.cap >= 4096)
    throw Error(CCP_ERR_STATE, "cell_capacity exceeded; raise cell_capacity");
  hip_check(hipStreamSynchronize(stream));
  drop_graphs();
  for (HostGrid *G : {&gf, &gh, &gx}) {
    G->g.cap *= 2;
    if (G->allocated) {
      G->bucket.free_();
      G->bucket.alloc((size_t)G->g.n_c * G->g.cap);
    }
    G->valid = false;
  }
  lg.cap = gf.g.cap;
  kp.gf = gf.g;
  kp.gh = gh.g;
  if (!split) kp.gh.n_c = 0;
  lgrid_valid = false;
}

// Cell lists of the LBS sites and of the current positions (fine grid of
// every bead, coarse grid of het beads); grows the capacity until every
// cell fits (not used inside the step loop).
void ccp_sim::ensure_grids() {
  if (gf.valid && (gh.valid || !split) && lgrid_valid) return;
  for (;;) {
    if (!lgrid_valid) build_grid(gf, lg, lr.p, (int)n_l, false, lspos.p, lstart.p, nullptr);
    if (!gf.valid) build_grid(gf, gf.g, r.p, (int)N, false, gf.spos.p, gf.cstart.p, nullptr, typed);
    if (split && !gh.valid) build_grid(gh, gh.g, r.p, (int)N, true, gh.spos.p, gh.cstart.p, nullptr);
    if (!take_overflow()) break;
    grow_capacity();
  }
  gf.valid = lgrid_valid = true;
  gh.valid = split;
}

void ccp_sim::sync_step() {
  if (dstep_dirty) {
    hip_check(hipMemcpyAsync(dstep.p, &step, sizeof(uint64_t), hipMemcpyHostToDevice, stream));
    hip_check(hipStreamSynchronize(stream));
    dstep_dirty = false;
  }
}

void ccp_sim::launch_stage(int mode, int stage, hipEvent_t mid) {
  KParams k = mode_params(mode);
  const bool het = split && mode != MODE_ICG;
  if (!het) k.gh.n_c = 0; // no het grid (single grid, or the soft relaxation)
  KPtrs q{};
  q.fspos = gf.spos.p;
  q.fcstart = gf.cstart.p;
  q.hspos = gh.spos.p;
  q.hcstart = gh.cstart.p;
  q.lspos = lspos.p;
  q.lstart = lstart.p;
  q.pos = stage == 1 ? r.p : er.p;
  q.out0 = stage == 1 ? er.p : r.p;
  q.pstar = pstar.p;
  q.fcount = gf.count.p;
  q.fblocksum = gf.bs(gf.clean);
  q.fhblocksum = typed ? gf.hbs(gf.clean) : nullptr;
  q.work = typed ? gf.work.p : nullptr;
  q.fbucket = gf.bucket.p;
  q.hcount = gh.count.p;
  q.hblocksum = gh.bs(gh.clean);
  q.hbucket = gh.bucket.p;
  q.flags = flags.p;
  q.step = dstep.p;
  q.fpair = fpair.p;
  q.noise = znoise.p;
  q.fbond = bond_pass() ? fbond.p : nullptr;
  // The next step's normals (the step counter still holds this step until
  // the order pass after stage 2 advances it): drawn in the tail of the
  // second stage's force launch by the type-ordered kernel, which does not
  // read them, otherwise by the order pass after the first stage.
  const NoiseJob nj = noise_job(1);
  const bool drew = launch_force(stream, mode, stage, k, q, typed && stage == 2 ? &nj : nullptr);
  if (typed && stage == 2 && !drew) throw Error(CCP_ERR_STATE, "type-ordered force launch without its normals");
  if (mid) hip_check(hipEventRecord(mid, stream));
  const OrderJob jf = gf.order_job(gf.g, typed, (uint32_t)kp.n_het, N, gf.spos.p, gf.cstart.p, nullptr);
  OrderJob jh{gh.g, gh.count.p, gh.bs(gh.clean), gh.bs(gh.clean ^ 1), gh.bucket.p, gh.spos.p, nullptr,
              gh.cstart.p};
  // and with the order pass, the bonded terms of q.out0
  const BondJob bj = bond_job(q.out0, mode);
  launch_order(stream, jf, het ? &jh : nullptr, flags.p, stage == 2 ? dstep.p : nullptr,
               stage == 1 && !typed ? &nj : nullptr, &k, bond_pass() ? &bj : nullptr);
  gf.clean ^= 1;
  if (het) gh.clean ^= 1;
  else gh.valid = false;
  gx.valid = false;
}

// The captured block of `len` steps for the current buffer parity (captured
// on first use).
hipGraphExec_t ccp_sim::step_graph(uint32_t len, int mode) {
  hipGraphExec_t &ge = graphs[graph_key(mode, gf.clean, gh.clean, len)];
  if (!ge) {
    const int cf = gf.clean, ch = gh.clean;
    const bool vh = gh.valid;
    hipGraph_t g;
    hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    for (uint32_t s = 0; s < len; ++s) launch_step(mode);
    hip_check(hipStreamEndCapture(stream, &g));
    hip_check(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    hip_check(hipGraphDestroy(g));
    // its device-side setup now, not at its first launch
    hip_check(hipGraphUpload(ge, stream));
    // each step flips the clean buffers twice
    gf.clean = cf;
    gh.clean = ch;
    gh.valid = mode != MODE_ICG ? vh : false;
  }
  return ge;
}

// n steps without contact sampling: the rest below graph_steps as graphs of
// the powers of two in its binary expansion, then whole graphs of
// graph_steps.
void ccp_sim::run_chunk(uint64_t n, int mode) {
  if (n == 0) return;
  if (!use_graphs) {
    for (uint64_t s = 0; s < n; ++s) launch_step(mode);
    hip_check(hipGetLastError());
    return;
  }
  const uint64_t G = graph_steps;
  const int cf = gf.clean, ch = gh.clean;
  const bool vh = gh.valid;
  // the blocks below G first, shortest first: the device starts on the first
  // block while the host still submits the longer ones
  const uint64_t full = n / G, rest = n % G;
  for (uint32_t len = 1; len <= (uint32_t)rest; len <<= 1)
    if (rest & len) hip_check(hipGraphLaunch(step_graph(len, mode), stream));
  if (full) {
    hipGraphExec_t ge = step_graph((uint32_t)G, mode);
    for (uint64_t b = 0; b < full; ++b) hip_check(hipGraphLaunch(ge, stream));
  }
  gf.clean = cf;
  gh.clean = ch;
  gh.valid = mode != MODE_ICG ? vh : false;
  hip_check(hipGetLastError());
}

void ccp_sim::capture_graphs(int mode) {
  if (!use_graphs) return;
  for (uint32_t len = 1; len < graph_steps; len <<= 1) step_graph(len, mode);
  step_graph(graph_steps, mode);
}

void ccp_sim::run_steps(uint64_t n, int mode) {
  ensure_grids();
  prime_terms(mode);
  while (n > 0) {
    uint64_t chunk = n;
    if (cinterval && mode != MODE_ICG) chunk = std::min<uint64_t>(n, cinterval - step % cinterval);
    run_chunk(chunk, mode);
    step += chunk;
    noise_step = step;
    n -= chunk;
    if (cinterval && mode != MODE_ICG && step % cinterval == 0) {
      launch_contacts(stream, kp, gf.spos.p, gf.cstart.p, cmap.p);
      ++n_samples;
    }
  }
}

// ---------------------------------------------------------------- initial condition
// chrsim::set_lbs_positions (chrsim.cu:596-631): uniform directions on the
// sphere of radius R_n - rco, rejecting sites closer than 2 lco to another
// or inside the opening; a hash grid of cell 2 lco replaces the O(n_l^2) scan.
void ccp_sim::set_lbs_positions(HostRng &g) {
  if (n_l == 0) return;
  const float hc = 2.0f * LCO;
  const int hs = 2 * (int)std::ceil(R_n / hc) + 3;
  std::map<long long, std::vector<uint32_t>> hash;
  auto key = [&](V3 p, int dx, int dy, int dz) {
    long long ix = (long long)std::floor(p.x / hc) + hs / 2 + dx, iy = (long long)std::floor(p.y / hc) + hs / 2 + dy,
              iz = (long long)std::floor(p.z / hc) + hs / 2 + dz;
    return (iz * hs + iy) * hs + ix;
  };
  uint64_t attempts = 0;
  for (uint32_t i_l = 0; i_l < n_l; ++i_l) {
    if (++attempts > 200000000ull) throw Error(CCP_ERR_PARAM, "unable to place lamina binding sites");
    float ran = g.uniform();
    float theta = (float)std::acos(1.0 - 2.0 * ran);
    ran = g.uniform();
    float phi = (float)(2.0 * M_PI * ran);
    V3 dir = {std::sin(theta) * std::cos(phi), std::sin(theta) * std::sin(phi), std::cos(theta)};
    V3 p = (R_n - RCO) * dir;
    bool ok = true;
    for (int dz = -1; dz <= 1 && ok; ++dz)
      for (int dy = -1; dy <= 1 && ok; ++dy)
        for (int dx = -1; dx <= 1 && ok; ++dx) {
          auto it = hash.find(key(p, dx, dy, dz));
          if (it == hash.end()) continue;
          for (uint32_t j : it->second) {
            V3 q = {hlr[3 * j], hlr[3 * j + 1], hlr[3 * j + 2]};
            if (length(q - p) < 2.0f * LCO) { ok = false; break; }
          }
        }
    if (p.z / length(p) > noc) ok = false;
    if (!ok) { --i_l; continue; }
    hlr[3 * i_l] = p.x; hlr[3 * i_l + 1] = p.y; hlr[3 * i_l + 2] = p.z;
    hash[key(p, 0, 0, 0)].push_back(i_l);
  }
}

// chrsim::set_particle_types (chrsim.cu:634-675) in ccp mode; het/LAD
// generator otherwise (het blocks Geometric(het_block_mean) at het_fraction,
// LAD blocks placed with an end-weighted profile).
void ccp_sim::set_particle_types(HostRng &g) {
  if (seq_mode == 0) {
    if (N == N_DEF) {
      std::ifstream f(seq_file);
      if (!f.is_open()) throw Error(CCP_ERR_IO, "unable to open " + seq_file);
      char c;
      for (uint32_t i = 0; i < N; ++i) {
        f >> c;
        hfl[i] = (c == 'A') ? (CCP_HET | CCP_LAD) : 0;
      }
      if (f.fail()) throw Error(CCP_ERR_IO, "failed to read sequence");
    } else {
      uint32_t cde = 0;
      bool a = false; // current type LNDe
      for (uint32_t i = 0; i < N; ++i) {
        if (i == cde) {
          float edl;
          do {
            float ran = g.uniform();
            edl = -MDL * std::log(ran);
          } while ((edl / MDL) > 5 || edl < 1.0f);
          cde = (uint32_t)(i + edl);
          a = !a;
        }
        hfl[i] = a ? (CCP_HET | CCP_LAD) : 0;
      }
    }
    return;
  }
  auto geometric = [&](double mean) {
    double p = 1.0 / mean;
    if (p >= 1.0) return (uint32_t)1;
    return 1u + (uint32_t)std::floor(std::log((double)g.uniform()) / std::log(1.0 - p));
  };
  const double f = het_fraction, m_h = het_mean, m_e = het_mean * (1.0 - f) / f;
  for (size_t c = 0; c + 1 < chain_off.size(); ++c) {
    const uint32_t s = chain_off[c], e = chain_off[c + 1], L = e - s;
    bool het = g.uniform() < (float)f;
    uint32_t i = s;
    while (i < e) {
      uint32_t len = geometric(het ? m_h : m_e);
      for (uint32_t k = 0; k < len && i < e; ++k, ++i) hfl[i] = het ? CCP_HET : 0;
      het = !het;
    }
    // LAD blocks
    const uint32_t target = (uint32_t)std::lround(lad_fraction * L);
    const double lam = std::max(1.0, lad_fraction * L / 2.0);
    uint32_t marked = 0;
    uint64_t guard = 0;
    while (marked < target && ++guard < 100ull * L + 1000) {
      uint32_t pos;
      if (lad_ends) {
        double d = -lam * std::log((double)g.uniform());
        if (d >= L) continue;
        uint32_t off = (uint32_t)d;
        pos = g.uniform() < 0.5f ? s + off : e - 1 - off;
      } else {
        pos = s + std::min<uint32_t>(L - 1, (uint32_t)(g.uniform() * L));
      }
      uint32_t len = geometric(m_h);
      for (uint32_t k = 0; k < len && marked < target; ++k) {
        uint32_t q = pos + k;
        if (q >= e) break;
        if (!(hfl[q] & CCP_LAD)) {
          hfl[q] |= CCP_LAD;
          ++marked;
        }
      }
    }
  }
}

// chrsim::perform_random_walk (chrsim.cu:719-802): confined random walk with
// bending-distributed bond angles, Gaussian bond lengths and ccp's
// backtracking rule.
void ccp_sim::perform_random_walk(HostRng &g, uint32_t i_s, uint32_t i_e, V3 dir, float mda) {
  const float len_d = length(dir);
  const bool cone = mda < (float)M_PI && len_d > 0.0f;
  const double iT = 1.0 / ((double)K_B * T);
  const double bend_scale = (ff == CCP_FF_LJ) ? 0.5 * k_b : 0.5; // ccp: 0.5 iT for k_b = 1
  const double ke = K_E;
  auto rand_dir = [&]() {
    float ran = g.uniform();
    float theta = (float)std::acos(1.0 - 2.0 * ran);
    ran = g.uniform();
    float phi = (float)(2.0 * M_PI * ran);
    return V3{std::sin(theta) * std::cos(phi), std::sin(theta) * std::sin(phi), std::cos(theta)};
  };
  auto in_cone = [&](V3 p, float d_r) { return !cone || !(std::acos(dot(p, dir) / (d_r * len_d)) > mda); };
  V3 old_dir;
  uint64_t guard = 0;
  for (;;) {
    V3 rd = rand_dir();
    float ran = g.uniform();
    float d_r = (float)((R_n - MIS) * std::pow((double)ran, 1.0 / 3));
    V3 p = d_r * rd;
    hr[3 * i_s] = p.x; hr[3 * i_s + 1] = p.y; hr[3 * i_s + 2] = p.z;
    old_dir = rd;
    if (in_cone(p, d_r)) break;
    if (++guard > 100000000ull) throw Error(CCP_ERR_PARAM, "unable to start random walk");
  }
  uint32_t att = 0;
  guard = 0;
  for (uint32_t i_p = i_s + 1; i_p < i_e; ++i_p) {
    if (++guard > 4000000000ull) throw Error(CCP_ERR_PARAM, "unable to complete random walk");
    V3 rd = rand_dir();
    V3 per = cross(old_dir, rd);
    per = (1.0f / std::sqrt(dot(per, per))) * per;
    float ran = g.uniform();
    const double beta = bend_scale * iT;
    float angle_b = (float)std::acos(1.0 + std::log(1.0 - (1.0 - std::exp(-2.0 * beta)) * ran) / beta);
    V3 nd = std::cos(angle_b) * old_dir + std::sin(angle_b) * per;
    ran = g.uniform();
    float len_b = (float)(L_0 + std::sqrt(2.0 / (ke * iT)) * erfinv(2.0 * ran - 1.0));
    V3 prev = {hr[3 * (i_p - 1)], hr[3 * (i_p - 1) + 1], hr[3 * (i_p - 1) + 2]};
    V3 p = len_b * nd + prev;
    hr[3 * i_p] = p.x; hr[3 * i_p + 1] = p.y; hr[3 * i_p + 2] = p.z;
    bool ok = std::isfinite(p.x) && std::isfinite(p.y) && std::isfinite(p.z);
    float d_r = length(p);
    if ((R_n - d_r) < MIS) ok = false;
    if (!in_cone(p, d_r)) ok = false;
    if (ok) {
      att = 0;
      old_dir = nd;
    } else {
      ++att;
      if (att > 1024) i_p = i_s + 1 + (i_p - i_s) * 3 / 4;
      else --i_p;
    }
  }
}

// chrsim::set_particle_positons (chrsim.cu:678-716): one walk per chain;
// several chains (ccp: the 12 chromosomes of N = 30386) start inside cones
// around shuffled icosahedron directions.
void ccp_sim::set_particle_positions(HostRng &g) {
  const size_t nch = chain_off.size() - 1;
  if (nch > 1) {
    const float phi = (float)((1.0 + std::sqrt(5.0)) / 2.0);
    const float mda = (float)(M_PI / 2.0 - std::atan(phi));
    V3 dirs[12] = {{0, +1, +phi}, {+1, +phi, 0}, {+phi, 0, +1}, {0, +1, -phi}, {+1, -phi, 0}, {-phi, 0, +1},
                   {0, -1, +phi}, {-1, +phi, 0}, {+phi, 0, -1}, {0, -1, -phi}, {-1, -phi, 0}, {-phi, 0, -1}};
    for (uint32_t i_d = 11; i_d > 0; --i_d) {
      float ran = g.uniform();
      uint32_t i_r = (uint32_t)((i_d + 1) * (1.0 - ran));
      std::swap(dirs[i_d], dirs[i_r]);
    }
    for (size_t c = 0; c < nch; ++c) perform_random_walk(g, chain_off[c], chain_off[c + 1], dirs[c % 12], mda);
  } else {
    perform_random_walk(g, 0, N, {0, 0, 0}, (float)M_PI);
  }
}

uint64_t ccp_sim::count_overlaps(float dmin) {
  ensure_grids();
  hip_check(hipMemsetAsync(dcount.p, 0, sizeof(unsigned long long), stream));
  launch_overlaps(stream, kp, gf.spos.p, gf.cstart.p, 0u, dmin * dmin, dcount.p);
  unsigned long long c = 0;
  hip_check(hipMemcpyAsync(&c, dcount.p, sizeof(c), hipMemcpyDeviceToHost, stream));
  hip_check(hipStreamSynchronize(stream));
  return c;
}

// chrsim::generate_initial_condition (chrsim.cu:470-522)
void ccp_sim::generate_initial_condition() {
  HostRng g(seed);
  set_lbs_positions(g);
  set_particle_types(g);
  if (replica) {
    // the same sites and types as replica 0, an independent walk
    HostRng w(seed, replica);
    set_particle_positions(w);
  } else {
    set_particle_positions(g);
  }
  upload_lbs();
  upload_positions();
  uint64_t po = count_overlaps(MIS), ipo = po;
  uint32_t rounds = 0;
  while (po > 0) {
    progress((float)(100.0 * (ipo - po) / ipo));
    if (++rounds > 10000) throw Error(CCP_ERR_STATE, "unable to remove bead overlaps");
    // a round in which a cell overflowed (clumps of the raw walk) is undone
    // and redone with wider cells: the result does not depend on the cells'
    // storage (the batch engine relaxes many systems with shared cells)
    hip_check(hipMemcpyAsync(aux.p, r.p, (size_t)N * sizeof(float4), hipMemcpyDeviceToDevice, stream));
    for (;;) {
      run_steps(spf, MODE_ICG);
      if (!take_overflow()) break;
      hip_check(hipMemcpyAsync(r.p, aux.p, (size_t)N * sizeof(float4), hipMemcpyDeviceToDevice, stream));
      terms_mode = -1;
      step -= spf;
      dstep_dirty = true;
      grow_capacity();
    }
    check_flags();
    po = count_overlaps(MIS);
  }
  download_positions();
  step = 0;
  dstep_dirty = true;
  i_f = 0;
  t = 0.0f;
  record("initial condition generated");
}

void ccp_sim::write_frame_bin(std::ofstream &f) {
  // minimal .trr frame (chrdat.cu:168-201)
  uint32_t tb;
  std::memcpy(&tb, &t, 4);
  uint32_t header[18] = {1993, 1, 0, 0, 0, 0, 0, 0, 0, 0, 3 * N * 4, 0, 0, N, i_f, 0, tb, 0};
  f.write(reinterpret_cast<const char *>(header), sizeof(header));
  f.write(reinterpret_cast<const char *>(hr.data()), 12 * (size_t)N);
  if (f.fail()) throw Error(CCP_ERR_IO, "failed to write frame to binary file");
}

// ---------------------------------------------------------------- C ABI
namespace {
thread_local std::string g_create_error;

template <class F> int guarded(ccp_sim *h, F &&fn) {
  if (!h) {
    g_create_error = "null handle";
    return CCP_ERR_ARG;
  }
  try {
    hip_check(hipSetDevice(h->dev));
    fn();
    return CCP_OK;
  } catch (const Error &e) {
    h->err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    h->err = e.what();
    return CCP_ERR_ARG;
  }
}

void need(bool ok, const char *what) {
  if (!ok) throw Error(CCP_ERR_ARG, what);
}

const char *CKPT_MAGIC = "CCPMI001";
} // namespace

extern "C" {

ccp_sim *ccp_create(const char *param_text) { return ccp_create_logged(param_text, nullptr, nullptr); }

ccp_sim *ccp_create_logged(const char *param_text, ccp_log_fn fn, void *user) {
  std::unique_ptr<ccp_sim> h(new ccp_sim());
  h->log_fn = fn;
  h->log_user = user;
  try {
    ParMap par(param_text);
    h->configure(par);
    h->fill_kparams();
    h->allocate();
    h->upload_positions();
    h->upload_lbs();
    return h.release();
  } catch (const std::exception &e) {
    g_create_error = e.what();
    return nullptr;
  }
}

void ccp_destroy(ccp_sim *h) { delete h; }

const char *ccp_last_error(const ccp_sim *h) { return h ? h->err.c_str() : g_create_error.c_str(); }
} // extern "C"

void ccpmi::host::set_null_error(const std::string &msg) { g_create_error = msg; }

extern "C" {

int ccp_set_logger(ccp_sim *h, ccp_log_fn fn, void *user) {
  return guarded(h, [&] {
    h->log_fn = fn;
    h->log_user = user;
  });
}

int ccp_get_info(const ccp_sim *h, ccp_info *o) {
  if (!h || !o) return CCP_ERR_ARG;
  o->N = h->N;
  o->n_l = h->n_l;
  o->n_chains = (uint32_t)h->chain_off.size() - 1;
  o->force_field = (uint32_t)h->ff;
  o->cps = (uint32_t)h->gx.g.cps;
  o->n_c = (uint32_t)h->gx.g.n_c;
  o->cell_capacity = (uint32_t)h->gh.g.cap;
  o->fine_cps = (uint32_t)h->gf.g.cps;
  o->fine_csl = h->gf.g.csl;
  o->spf = h->spf;
  o->fpf = h->fpf;
  o->contact_bins = h->cbins;
  o->contact_bin_size = h->cbin_size;
  o->contact_interval = h->cinterval;
  o->csl = h->csl;
  o->R_n = h->R_n;
  o->R_o = h->R_o;
  o->R_b = h->R_b;
  o->T = h->T;
  o->dt = h->dt;
  o->sd = h->sd;
  o->contact_cutoff = h->contact_rc;
  o->force_cap = h->fcap;
  o->seed = h->seed;
  return CCP_OK;
}

int ccp_generate_initial_condition(ccp_sim *h) {
  return guarded(h, [&] { h->generate_initial_condition(); });
}

int ccp_save_checkpoint(ccp_sim *h, const char *path) {
  return guarded(h, [&] {
    need(path != nullptr, "null path");
    std::ofstream f(path, std::ios::binary);
    if (!f.is_open()) throw Error(CCP_ERR_IO, std::string("unable to open ") + path);
    h->download_positions();
    f.write(reinterpret_cast<const char *>(&h->i_f), 4);
    f.write(reinterpret_cast<const char *>(&h->t), 4);
    for (uint32_t i = 0; i < h->N; ++i) {
      int32_t pt = (h->hfl[i] & CCP_HET) ? 0 : 1;
      f.write(reinterpret_cast<const char *>(&pt), 4);
    }
    f.write(reinterpret_cast<const char *>(h->hr.data()), 12 * (size_t)h->N);
    f.write(reinterpret_cast<const char *>(h->hlr.data()), 12 * (size_t)h->n_l);
    f.write(CKPT_MAGIC, 8);
    f.write(reinterpret_cast<const char *>(&h->seed), 8);
    f.write(reinterpret_cast<const char *>(&h->step), 8);
    f.write(reinterpret_cast<const char *>(&h->N), 4);
    f.write(reinterpret_cast<const char *>(h->hfl.data()), h->N);
    if (f.fail()) throw Error(CCP_ERR_IO, "failed to write checkpoint");
    h->record("simulation checkpoint saved");
  });
}

int ccp_load_checkpoint(ccp_sim *h, const char *path) {
  return guarded(h, [&] {
    need(path != nullptr, "null path");
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) throw Error(CCP_ERR_IO, std::string("unable to open ") + path);
    uint32_t i_f;
    float t;
    std::vector<int32_t> pt(h->N);
    std::vector<float> r(3 * (size_t)h->N), lr(3 * (size_t)h->n_l);
    f.read(reinterpret_cast<char *>(&i_f), 4);
    f.read(reinterpret_cast<char *>(&t), 4);
    f.read(reinterpret_cast<char *>(pt.data()), 4 * (size_t)h->N);
    f.read(reinterpret_cast<char *>(r.data()), 12 * (size_t)h->N);
    f.read(reinterpret_cast<char *>(lr.data()), 12 * (size_t)h->n_l);
    if (f.fail()) throw Error(CCP_ERR_IO, "failed to read checkpoint");
    h->i_f = i_f;
    h->t = t;
    h->hr = r;
    h->hlr = lr;
    for (uint32_t i = 0; i < h->N; ++i) h->hfl[i] = pt[i] == 0 ? (CCP_HET | CCP_LAD) : 0;
    h->step = (uint64_t)i_f * h->spf;
    char magic[8];
    f.read(magic, 8);
    if (f && std::memcmp(magic, CKPT_MAGIC, 8) == 0) {
      uint64_t seed, step;
      uint32_t n;
      f.read(reinterpret_cast<char *>(&seed), 8);
      f.read(reinterpret_cast<char *>(&step), 8);
      f.read(reinterpret_cast<char *>(&n), 4);
      if (f && n == h->N) {
        std::vector<uint8_t> fl(n);
        f.read(reinterpret_cast<char *>(fl.data()), n);
        if (f) {
          h->hfl = fl;
          h->seed = seed;
          h->step = step;
          h->fill_kparams();
        }
      }
    }
    // Captured graphs hold the kernel parameters (the noise key among them)
    // by value, and znoise may hold normals of the old key for this step:
    // both are rebuilt from the loaded state.
    h->drop_graphs();
    h->noise_step = ~0ull;
    h->dstep_dirty = true;
    h->upload_positions();
    h->upload_lbs();
    h->record("simulation checkpoint loaded");
  });
}

int ccp_write_frame_txt(ccp_sim *h, const char *path) {
  return guarded(h, [&] {
    need(path != nullptr, "null path");
    std::ofstream f(path);
    if (!f.is_open()) throw Error(CCP_ERR_IO, std::string("unable to open ") + path);
    h->download_positions();
    // chrdat::write_frame_txt (chrdat.cu:107-133)
    f << h->i_f << " " << h->t << "\n";
    f << cnfs(h->N, 5, ' ') << "\n";
    for (uint32_t i = 0; i < h->N; ++i) {
      char ptc = (h->hfl[i] & CCP_HET) ? 'A' : 'B';
      f << std::setw(5) << i + 1 << std::left << std::setw(5) << ptc;
      f << std::right << std::setw(5) << ptc << std::setw(5) << i + 1;
      f << cnfs(h->hr[3 * i], 8, ' ', 3) << cnfs(h->hr[3 * i + 1], 8, ' ', 3) << cnfs(h->hr[3 * i + 2], 8, ' ', 3);
      f << "\n";
    }
    f << cnfs(0.0, 10, ' ', 5) << cnfs(0.0, 10, ' ', 5) << cnfs(0.0, 10, ' ', 5) << "\n";
    if (f.fail()) throw Error(CCP_ERR_IO, "failed to write frame to text file");
  });
}

int ccp_read_frame_txt(ccp_sim *h, const char *path) {
  return guarded(h, [&] {
    need(path != nullptr, "null path");
    std::ifstream f(path);
    if (!f.is_open()) throw Error(CCP_ERR_IO, std::string("unable to open ") + path);
    // chrdat::read_frame_txt (chrdat.cu:136-165)
    std::string aux;
    char ptc;
    uint32_t i_f;
    float t;
    f >> i_f >> t >> aux;
    std::vector<float> r(3 * (size_t)h->N);
    std::vector<uint8_t> fl(h->N);
    for (uint32_t i = 0; i < h->N; ++i) {
      f >> aux >> ptc >> aux;
      fl[i] = (ptc == 'A') ? (CCP_HET | CCP_LAD) : 0;
      f >> r[3 * i] >> r[3 * i + 1] >> r[3 * i + 2];
    }
    f >> aux >> aux >> aux;
    if (f.fail()) throw Error(CCP_ERR_IO, "failed to read frame from text file");
    // keep the LAD bits of a het/LAD sequence when the het pattern is unchanged
    for (uint32_t i = 0; i < h->N; ++i)
      if (h->ff == CCP_FF_LJ && ((h->hfl[i] & CCP_HET) != 0) == ((fl[i] & CCP_HET) != 0)) fl[i] = h->hfl[i];
    h->i_f = i_f;
    h->t = t;
    h->hr = r;
    h->hfl = fl;
    h->upload_positions();
  });
}

int ccp_write_lbs_txt(ccp_sim *h, const char *path) {
  return guarded(h, [&] {
    need(path != nullptr, "null path");
    std::ofstream f(path);
    if (!f.is_open()) throw Error(CCP_ERR_IO, std::string("unable to open ") + path);
    // chrdat::write_lbs_txt (chrdat.cu:229-254)
    f << "\n";
    f << cnfs(h->n_l, 5, ' ') << "\n";
    const char ptc = 'C';
    for (uint32_t l = 0; l < h->n_l; ++l) {
      f << std::setw(5) << l + 1 << std::left << std::setw(5) << ptc;
      f << std::right << std::setw(5) << ptc << std::setw(5) << l + 1;
      f << cnfs(h->hlr[3 * l], 8, ' ', 3) << cnfs(h->hlr[3 * l + 1], 8, ' ', 3) << cnfs(h->hlr[3 * l + 2], 8, ' ', 3);
      f << "\n";
    }
    f << cnfs(0.0, 10, ' ', 5) << cnfs(0.0, 10, ' ', 5) << cnfs(0.0, 10, ' ', 5) << "\n";
    if (f.fail()) throw Error(CCP_ERR_IO, "failed to write lamina binding sites to text file");
  });
}

int ccp_read_lbs_txt(ccp_sim *h, const char *path) {
  return guarded(h, [&] {
    need(path != nullptr, "null path");
    std::ifstream f(path);
    if (!f.is_open()) throw Error(CCP_ERR_IO, std::string("unable to open ") + path);
    // The file starts with an empty title line, then the site count
    // (chrdat.cu:229-236); ccp's reader (chrdat.cu:257-284) is never called
    // and skips one token too many, so the layout is read as written.
    std::string line, aux;
    char ptc;
    std::getline(f, line);
    f >> aux;
    std::vector<float> lr(3 * (size_t)h->n_l);
    for (uint32_t l = 0; l < h->n_l; ++l) {
      f >> aux >> ptc >> aux;
      f >> lr[3 * l] >> lr[3 * l + 1] >> lr[3 * l + 2];
    }
    f >> aux >> aux >> aux;
    if (f.fail()) throw Error(CCP_ERR_IO, "failed to read lamina binding sites from text file");
    h->hlr = lr;
    h->upload_lbs();
  });
}

int ccp_run_simulation(ccp_sim *h, const char *trr_path) {
  return guarded(h, [&] {
    std::ofstream f;
    if (trr_path) {
      f.open(trr_path, std::ios::binary);
      if (!f.is_open()) throw Error(CCP_ERR_IO, std::string("unable to open ") + trr_path);
    }
    // chrsim::run_simulation (chrsim.cu:558-593)
    h->lgrid_valid = false;
    for (uint32_t ffi = 0; ffi < h->fpf; ++ffi) {
      h->progress((float)(100.0 * ffi / h->fpf));
      h->run_steps(h->spf, h->ff == CCP_FF_LJ ? MODE_LJ : MODE_CCP);
      h->download_positions();
      h->check_flags();
      ++h->i_f;
      h->t += h->spf * h->dt;
      if (trr_path) h->write_frame_bin(f);
    }
    h->record("simulation ended");
  });
}

int ccp_run_frames(ccp_sim *h, uint32_t n_frames, ccp_frame_fn cb, void *user) {
  return guarded(h, [&] {
    need(cb != nullptr, "frame callback missing");
    h->lgrid_valid = false;
    for (uint32_t k = 0; k < n_frames; ++k) {
      h->run_steps(h->spf, h->ff == CCP_FF_LJ ? MODE_LJ : MODE_CCP);
      h->download_positions();
      h->check_flags();
      ++h->i_f;
      h->t += h->spf * h->dt;
      if (cb(h->i_f, h->t, h->hr.data(), h->N, user)) break;
    }
  });
}

int ccp_run_steps(ccp_sim *h, uint64_t n) {
  return guarded(h, [&] { h->run_steps(n, h->ff == CCP_FF_LJ ? MODE_LJ : MODE_CCP); });
}

int ccp_synchronize(ccp_sim *h) {
  return guarded(h, [&] { h->check_flags(); });
}

int ccp_get_time(const ccp_sim *h, uint32_t *i_f, float *t, uint64_t *step) {
  if (!h) return CCP_ERR_ARG;
  if (i_f) *i_f = h->i_f;
  if (t) *t = h->t;
  if (step) *step = h->step;
  return CCP_OK;
}

int ccp_set_time(ccp_sim *h, uint32_t i_f, float t, uint64_t step) {
  return guarded(h, [&] {
    h->i_f = i_f;
    h->t = t;
    h->step = step;
    h->dstep_dirty = true;
  });
}

int ccp_get_positions(ccp_sim *h, float *xyz, size_t n) {
  return guarded(h, [&] {
    need(xyz && n == h->N, "positions size mismatch");
    h->download_positions();
    std::memcpy(xyz, h->hr.data(), 12 * (size_t)h->N);
  });
}

int ccp_set_positions(ccp_sim *h, const float *xyz, size_t n) {
  return guarded(h, [&] {
    need(xyz && n == h->N, "positions size mismatch");
    std::memcpy(h->hr.data(), xyz, 12 * (size_t)h->N);
    h->upload_positions();
  });
}

int ccp_get_types(ccp_sim *h, int32_t *pt, size_t n) {
  return guarded(h, [&] {
    need(pt && n == h->N, "types size mismatch");
    for (uint32_t i = 0; i < h->N; ++i) pt[i] = (h->hfl[i] & CCP_HET) ? 0 : 1;
  });
}

int ccp_set_types(ccp_sim *h, const int32_t *pt, size_t n) {
  return guarded(h, [&] {
    need(pt && n == h->N, "types size mismatch");
    h->download_positions();
    for (uint32_t i = 0; i < h->N; ++i) h->hfl[i] = pt[i] == 0 ? (CCP_HET | CCP_LAD) : 0;
    h->upload_positions();
  });
}

int ccp_get_flags(ccp_sim *h, uint8_t *fl, size_t n) {
  return guarded(h, [&] {
    need(fl && n == h->N, "flags size mismatch");
    std::memcpy(fl, h->hfl.data(), h->N);
  });
}

int ccp_set_flags(ccp_sim *h, const uint8_t *fl, size_t n) {
  return guarded(h, [&] {
    need(fl && n == h->N, "flags size mismatch");
    h->download_positions();
    for (uint32_t i = 0; i < h->N; ++i) h->hfl[i] = fl[i] & (CCP_HET | CCP_LAD);
    h->upload_positions();
  });
}

int ccp_get_lbs(ccp_sim *h, float *xyz, size_t n_l) {
  return guarded(h, [&] {
    need(n_l == h->n_l && (xyz || n_l == 0), "lbs size mismatch");
    if (n_l) std::memcpy(xyz, h->hlr.data(), 12 * (size_t)h->n_l);
  });
}

int ccp_set_lbs(ccp_sim *h, const float *xyz, size_t n_l) {
  return guarded(h, [&] {
    need(n_l == h->n_l && (xyz || n_l == 0), "lbs size mismatch");
    if (n_l) std::memcpy(h->hlr.data(), xyz, 12 * (size_t)h->n_l);
    h->upload_lbs();
  });
}

int ccp_get_chains(ccp_sim *h, uint32_t *off, size_t n) {
  return guarded(h, [&] {
    need(off && n == h->chain_off.size(), "chains size mismatch");
    std::memcpy(off, h->chain_off.data(), 4 * n);
  });
}

int ccp_compute_forces(ccp_sim *h, int ff, float *f_xyz, size_t n) {
  return guarded(h, [&] {
    need(f_xyz && n == h->N, "forces size mismatch");
    need(ff == CCP_FF_CCP || ff == CCP_FF_ICG || ff == CCP_FF_LJ, "unknown force field");
    h->ensure_grids();
    KParams k = h->mode_params(ff);
    KPtrs q{};
    q.fspos = h->gf.spos.p;
    q.fcstart = h->gf.cstart.p;
    q.hspos = h->gh.spos.p;
    q.hcstart = h->gh.cstart.p;
    q.lspos = h->lspos.p;
    q.lstart = h->lstart.p;
    q.pos = h->r.p;
    q.out0 = h->aux.p;
    q.flags = h->flags.p;
    q.step = h->dstep.p;
    q.fpair = h->fpair.p;
    // the grid build above wrote the type-ordered work list: the forces come
    // from the kernel the steps run
    q.work = h->typed ? h->gf.work.p : nullptr;
    launch_force(h->stream, ff, 0, k, q);
    hip_check(hipGetLastError());
    std::vector<float4> out(h->N);
    hip_check(hipMemcpyAsync(out.data(), h->aux.p, h->N * sizeof(float4), hipMemcpyDeviceToHost, h->stream));
    hip_check(hipStreamSynchronize(h->stream));
    for (uint32_t i = 0; i < h->N; ++i) {
      f_xyz[3 * i] = out[i].x;
      f_xyz[3 * i + 1] = out[i].y;
      f_xyz[3 * i + 2] = out[i].z;
    }
  });
}

int ccp_cell_list(ccp_sim *h, uint32_t *sci, uint32_t *soi, uint32_t *beg, uint32_t *end, size_t n, size_t n_c) {
  return guarded(h, [&] {
    auto &G = h->gx;
    need(n == h->N && n_c == (size_t)G.g.n_c && sci && soi && beg && end, "cell list size mismatch");
    if (!G.allocated) G.alloc(h->N);
    for (;;) {
      h->build_grid(G, G.g, h->r.p, (int)h->N, false, G.spos.p, G.cstart.p, G.soi.p);
      if (!h->take_overflow()) break;
      h->grow_capacity();
    }
    G.valid = true;
    std::vector<uint32_t> cs((size_t)G.g.n_c + 1);
    hip_check(hipMemcpyAsync(soi, G.soi.p, 4 * n, hipMemcpyDeviceToHost, h->stream));
    hip_check(hipMemcpyAsync(cs.data(), G.cstart.p, 4 * (n_c + 1), hipMemcpyDeviceToHost, h->stream));
    hip_check(hipStreamSynchronize(h->stream));
    h->check_flags();
    // ccp layout: beg/end of non-empty cells, 0/0 for empty ones (sugrid.cu:38-77)
    for (size_t c = 0; c < n_c; ++c) {
      const bool full = cs[c + 1] > cs[c];
      beg[c] = full ? cs[c] : 0u;
      end[c] = full ? cs[c + 1] : 0u;
      for (uint32_t s = cs[c]; s < cs[c + 1]; ++s) sci[s] = (uint32_t)c;
    }
  });
}

int ccp_step_grid(ccp_sim *h, uint32_t *cstart, size_t n_c1, uint32_t *soi, size_t n) {
  return guarded(h, [&] {
    need(cstart && soi && n == h->N && n_c1 == (size_t)h->gf.g.n_c + 1, "step grid size mismatch");
    // the grid of the current positions: the one the last stage binned and
    // ordered, or a fresh build when the positions were set since
    h->ensure_grids();
    std::vector<float4> sp(h->N);
    hip_check(hipMemcpyAsync(cstart, h->gf.cstart.p, 4 * n_c1, hipMemcpyDeviceToHost, h->stream));
    hip_check(hipMemcpyAsync(sp.data(), h->gf.spos.p, sizeof(float4) * n, hipMemcpyDeviceToHost, h->stream));
    hip_check(hipStreamSynchronize(h->stream));
    h->check_flags();
    for (size_t k = 0; k < n; ++k) {
      uint32_t w;
      std::memcpy(&w, &sp[k].w, 4);
      soi[k] = w & ID_MASK;
    }
  });
}

int ccp_noise(ccp_sim *h, uint64_t step, float *z_xyz, uint32_t *attempts, size_t n) {
  return guarded(h, [&] {
    need(z_xyz && n == h->N, "noise size mismatch");
    launch_noise(h->stream, h->kp, step, h->aux.p);
    hip_check(hipGetLastError());
    std::vector<float4> out(h->N);
    hip_check(hipMemcpyAsync(out.data(), h->aux.p, h->N * sizeof(float4), hipMemcpyDeviceToHost, h->stream));
    hip_check(hipStreamSynchronize(h->stream));
    for (uint32_t i = 0; i < h->N; ++i) {
      z_xyz[3 * i] = out[i].x;
      z_xyz[3 * i + 1] = out[i].y;
      z_xyz[3 * i + 2] = out[i].z;
      if (attempts) std::memcpy(&attempts[i], &out[i].w, 4);
    }
  });
}

int ccp_count_overlaps(ccp_sim *h, float dmin, uint64_t *count) {
  return guarded(h, [&] {
    need(count != nullptr, "null count");
    need(dmin > 0.0f && dmin <= h->gf.g.csl, "dmin must be in (0, fine_cell_size]");
    *count = h->count_overlaps(dmin);
  });
}

int ccp_contact_sample(ccp_sim *h) {
  return guarded(h, [&] {
    h->ensure_grids();
    launch_contacts(h->stream, h->kp, h->gf.spos.p, h->gf.cstart.p, h->cmap.p);
    hip_check(hipGetLastError());
    ++h->n_samples;
  });
}

int ccp_contact_reset(ccp_sim *h) {
  return guarded(h, [&] {
    hip_check(hipMemsetAsync(h->cmap.p, 0, (size_t)h->cbins * h->cbins * 4, h->stream));
    hip_check(hipStreamSynchronize(h->stream));
    h->n_samples = 0;
  });
}

int ccp_contact_map(ccp_sim *h, uint32_t *map, size_t n_bins2, uint32_t *n_samples) {
  return guarded(h, [&] {
    need(map && n_bins2 == (size_t)h->cbins * h->cbins, "contact map size mismatch");
    hip_check(hipMemcpyAsync(map, h->cmap.p, 4 * n_bins2, hipMemcpyDeviceToHost, h->stream));
    hip_check(hipStreamSynchronize(h->stream));
    if (n_samples) *n_samples = h->n_samples;
  });
}

int ccp_observables(ccp_sim *h, double scalars[4], double *rcd, size_t n_bins) {
  return guarded(h, [&] {
    need(scalars && rcd && n_bins >= 1 && n_bins <= 4096, "observables size mismatch");
    const size_t n_acc = 6 + 3 * n_bins;
    DevBuf<double> acc;
    acc.alloc(n_acc);
    hip_check(hipMemsetAsync(acc.p, 0, n_acc * sizeof(double), h->stream));
    launch_observables(h->stream, (int)h->N, (double)h->R_n, (double)h->nod, (double)RCO / 2.0, (int)n_bins, h->r.p,
                       acc.p);
    hip_check(hipGetLastError());
    std::vector<double> a(n_acc);
    hip_check(hipMemcpyAsync(a.data(), acc.p, n_acc * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    hip_check(hipStreamSynchronize(h->stream));
    acc.free_();
    const double n = h->N, cx = a[0] / n, cy = a[1] / n, cz = a[2] / n;
    scalars[0] = std::sqrt(cx * cx + cy * cy + cz * cz);
    scalars[1] = a[3] / n - (cx * cx + cy * cy + cz * cz);
    scalars[2] = h->N > 1 ? a[4] / (n - 1.0) : 0.0;
    scalars[3] = a[5] / n;
    const double shell = 4.0 / 3.0 * M_PI * (double)h->R_n * h->R_n * h->R_n / (double)n_bins;
    for (size_t k = 0; k < 3 * n_bins; ++k) rcd[k] = a[6 + k] / shell;
  });
}

int ccp_distance_contact(ccp_sim *h, uint32_t chromosome, float *sd, float *cp, size_t n) {
  return guarded(h, [&] {
    uint32_t i_s = 0, i_e = 0;
    need(sd && cp && h->sd_cp_range(chromosome, i_s, i_e), "no such chromosome");
    const uint32_t lsdcp = (i_e - i_s) / 2;
    need(n == lsdcp, "distance/contact size mismatch");
    DevBuf<float> out;
    out.alloc(2 * (size_t)lsdcp);
    launch_sd_cp(h->stream, h->r.p, h->N, lsdcp, i_s, i_e, out.p, out.p + lsdcp, nullptr, nullptr);
    hip_check(hipGetLastError());
    hip_check(hipMemcpyAsync(sd, out.p, lsdcp * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    hip_check(hipMemcpyAsync(cp, out.p + lsdcp, lsdcp * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    hip_check(hipStreamSynchronize(h->stream));
    out.free_();
  });
}

int ccp_frame_contact_map(ccp_sim *h, float *cm, size_t lcm) {
  return guarded(h, [&] {
    const uint32_t cms = (h->N == N_DEF) ? (h->N / 2) / 4 : h->N / 4;
    need(cm && lcm == (size_t)cms * (cms + 1) / 2, "contact map size mismatch");
    DevBuf<float> out;
    out.alloc(lcm);
    launch_ccp_cm(h->stream, h->N, (uint32_t)lcm, h->r.p, out.p, nullptr);
    hip_check(hipGetLastError());
    hip_check(hipMemcpyAsync(cm, out.p, lcm * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    hip_check(hipStreamSynchronize(h->stream));
    out.free_();
  });
}

int ccp_time_kernels(ccp_sim *h, uint32_t n_steps, float out[5]) {
  return guarded(h, [&] {
    need(out && n_steps > 0 && n_steps <= 100000, "bad timing request");
    const int mode = h->ff == CCP_FF_LJ ? MODE_LJ : MODE_CCP;
    h->ensure_grids();
    h->prime_terms(mode);
    // Events before the force launch, between it and the order launch, and
    // after the order launch of every stage, all recorded in the stream with
    // no wait in between, so the kernels run back to back as in a run; the
    // host reads them once at the end.
    std::vector<hipEvent_t> ev(6 * (size_t)n_steps);
    for (auto &e : ev) hip_check(hipEventCreate(&e));
    size_t k = 0;
    for (uint32_t s = 0; s < n_steps; ++s)
      for (int stage = 1; stage <= 2; ++stage) {
        hip_check(hipEventRecord(ev[k], h->stream));
        h->launch_stage(mode, stage, ev[k + 1]);
        hip_check(hipEventRecord(ev[k + 2], h->stream));
        k += 3;
      }
    hip_check(hipEventSynchronize(ev.back()));
    double acc[4] = {0, 0, 0, 0};
    k = 0;
    for (uint32_t s = 0; s < n_steps; ++s)
      for (int stage = 1; stage <= 2; ++stage) {
        float a, b;
        hip_check(hipEventElapsedTime(&a, ev[k], ev[k + 1]));
        hip_check(hipEventElapsedTime(&b, ev[k + 1], ev[k + 2]));
        acc[(stage - 1) * 2] += a;
        acc[(stage - 1) * 2 + 1] += b;
        k += 3;
      }
    float whole;
    hip_check(hipEventElapsedTime(&whole, ev.front(), ev.back()));
    h->step += n_steps;
    h->noise_step = h->step;
    for (int j = 0; j < 4; ++j) out[j] = (float)(acc[j] / n_steps);
    out[4] = whole / n_steps;
    for (auto &e : ev) (void)hipEventDestroy(e);
    h->check_flags();
  });
}

int ccp_capture_graphs(ccp_sim *h) {
  return guarded(h, [&] {
    const int mode = h->ff == CCP_FF_LJ ? MODE_LJ : MODE_CCP;
    h->ensure_grids();
    h->capture_graphs(mode);
  });
}

// ---------------------------------------------------------------- analysis
} // extern "C"

// ccp's chrana state (chrana.cuh:56-90): time series of the frame
// observables on the host; the basic observables (d(s), P(s), contact map)
// as FP64 sums on the device, closed per simulation into sum avg and
// sum avg^2 for the combined statistics.
struct ccp_analysis {
  static constexpr uint32_t N_B = 64; // radial bins (chrana.cuh:12)
  static constexpr uint32_t N_OBS = 4 + 3 * N_B;
  ccp_sim *h = nullptr;
  int dev = 0;
  std::string err;
  uint32_t cms = 0, lcm = 0, n_s = 0, n_dp = 0;
  uint32_t lsdcp[6] = {}, i_s[6] = {}, i_e[6] = {};
  std::vector<float> ts[N_OBS];           // dcm, rg2, nop, ncf, rcd[3][64]
  std::vector<SeriesStat> sv[N_OBS];      // per simulation
  DevBuf<double> obs;                     // per-frame observable sums
  DevBuf<double> sdcp[6];                 // sum_sd, sum_cp, m1_sd, m1_cp, m2_sd, m2_cp (lsdcp each)
  DevBuf<double> cm;                      // sum, m1, m2 (lcm each)
  ~ccp_analysis() {
    obs.free_();
    for (auto &b : sdcp) b.free_();
    cm.free_();
  }
};

namespace {

template <class F> int aguarded(ccp_analysis *a, F &&fn) {
  if (!a) return CCP_ERR_ARG;
  try {
    hip_check(hipSetDevice(a->h->dev));
    fn();
    return CCP_OK;
  } catch (const Error &e) {
    a->err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    a->err = e.what();
    return CCP_ERR_ARG;
  }
}

// "avg sqrt(var) sem" columns of ccp's statistics lines (sostat.cu:21-47)
std::string stat3(const SeriesStat &s) {
  return cnfs(s.avg, 12, ' ', 6) + cnfs(std::sqrt(s.var), 12, ' ', 6) + cnfs(s.sem, 12, ' ', 6);
}

void add_frame(ccp_analysis *a) {
  ccp_sim *h = a->h;
  constexpr uint32_t nb = ccp_analysis::N_B;
  const size_t n_acc = 6 + 3 * nb;
  hip_check(hipMemsetAsync(a->obs.p, 0, n_acc * sizeof(double), h->stream));
  launch_observables(h->stream, (int)h->N, (double)h->R_n, (double)h->nod, (double)RCO / 2.0, (int)nb, h->r.p,
                     a->obs.p);
  for (uint32_t c = 0; c < 6; ++c) {
    const uint32_t L = a->lsdcp[c];
    if (!L) continue;
    double *b = a->sdcp[c].p;
    launch_sd_cp(h->stream, h->r.p, h->N, L, a->i_s[c], a->i_e[c], nullptr, nullptr, b, b + L);
  }
  launch_ccp_cm(h->stream, h->N, a->lcm, h->r.p, nullptr, a->cm.p);
  hip_check(hipGetLastError());
  std::vector<double> acc(n_acc);
  hip_check(hipMemcpyAsync(acc.data(), a->obs.p, n_acc * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  hip_check(hipStreamSynchronize(h->stream));
  // the series hold FP32 values, as ccp's (FP64 sums here, FP32 in ccp)
  const double n = h->N, cx = acc[0] / n, cy = acc[1] / n, cz = acc[2] / n;
  a->ts[0].push_back((float)std::sqrt(cx * cx + cy * cy + cz * cz));
  a->ts[1].push_back((float)(acc[3] / n - (cx * cx + cy * cy + cz * cz)));
  a->ts[2].push_back((float)(h->N > 1 ? acc[4] / (n - 1.0) : 0.0));
  a->ts[3].push_back((float)(acc[5] / n));
  const double shell = 4.0 / 3.0 * M_PI * (double)h->R_n * h->R_n * h->R_n / nb;
  for (uint32_t k = 0; k < 3 * nb; ++k) a->ts[4 + k].push_back((float)(acc[6 + k] / shell));
  ++a->n_dp;
}

} // namespace

extern "C" {

int ccp_analysis_create(ccp_sim *h, ccp_analysis **out) {
  return guarded(h, [&] {
    need(out != nullptr, "null output");
    *out = nullptr;
    std::unique_ptr<ccp_analysis> a(new ccp_analysis());
    a->h = h;
    a->dev = h->dev;
    // chrana::chrana (chrana.cu:88-120)
    a->cms = (h->N == N_DEF) ? (h->N / 2) / 4 : h->N / 4;
    need(a->cms >= 1 && a->cms <= 65535, "contact map side out of range");
    a->lcm = a->cms * (a->cms + 1) / 2;
    for (uint32_t c = 0; c < 6; ++c) {
      uint32_t b, e;
      if (h->sd_cp_range(c, b, e)) {
        a->i_s[c] = b;
        a->i_e[c] = e;
        a->lsdcp[c] = (e - b) / 2;
      }
      a->sdcp[c].alloc(6 * (size_t)a->lsdcp[c]);
      if (a->lsdcp[c]) hip_check(hipMemsetAsync(a->sdcp[c].p, 0, 6 * (size_t)a->lsdcp[c] * 8, h->stream));
    }
    a->obs.alloc(6 + 3 * ccp_analysis::N_B);
    a->cm.alloc(3 * (size_t)a->lcm);
    hip_check(hipMemsetAsync(a->cm.p, 0, 3 * (size_t)a->lcm * 8, h->stream));
    hip_check(hipStreamSynchronize(h->stream));
    *out = a.release();
  });
}

void ccp_analysis_destroy(ccp_analysis *a) {
  if (a) (void)hipSetDevice(a->dev);
  delete a;
}

const char *ccp_analysis_last_error(const ccp_analysis *a) { return a ? a->err.c_str() : "null handle"; }

int ccp_analysis_add_frame(ccp_analysis *a) {
  return aguarded(a, [&] { add_frame(a); });
}

int ccp_analysis_add_trajectory_file(ccp_analysis *a, const char *path) {
  return aguarded(a, [&] {
    need(path != nullptr, "null path");
    ccp_sim *h = a->h;
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) throw Error(CCP_ERR_IO, std::string("unable to open ") + path);
    for (uint32_t ffi = 0; ffi < h->fpf; ++ffi) {
      h->progress((float)(100.0 * ffi / h->fpf));
      // chrdat::read_frame_bin (chrdat.cu:204-226)
      uint32_t header[18];
      f.read(reinterpret_cast<char *>(header), sizeof(header));
      std::vector<float> r(3 * (size_t)h->N);
      f.read(reinterpret_cast<char *>(r.data()), 12 * (size_t)h->N);
      if (f.fail()) throw Error(CCP_ERR_IO, "failed to read frame from binary file");
      h->i_f = header[14];
      std::memcpy(&h->t, &header[16], 4);
      h->hr = r;
      h->upload_positions();
      add_frame(a);
    }
  });
}

int ccp_analysis_end_simulation(ccp_analysis *a, const char *path) {
  return aguarded(a, [&] {
    need(path != nullptr, "null path");
    need(a->n_dp > 0, "no frames added");
    ccp_sim *h = a->h;
    constexpr uint32_t nb = ccp_analysis::N_B;
    // calc_last_is_stat (chrana.cu:154-181)
    for (uint32_t k = 0; k < ccp_analysis::N_OBS; ++k)
      a->sv[k].push_back(series_stats(a->ts[k].data(), a->ts[k].size()));
    std::vector<std::vector<double>> sd(6), cp(6);
    for (uint32_t c = 0; c < 6; ++c) {
      const uint32_t L = a->lsdcp[c];
      sd[c].resize(L);
      cp[c].resize(L);
      if (!L) continue;
      hip_check(hipMemcpyAsync(sd[c].data(), a->sdcp[c].p, L * 8, hipMemcpyDeviceToHost, h->stream));
      hip_check(hipMemcpyAsync(cp[c].data(), a->sdcp[c].p + L, L * 8, hipMemcpyDeviceToHost, h->stream));
    }
    hip_check(hipStreamSynchronize(h->stream));
    // save_last_is_stat (chrana.cu:184-247)
    std::ofstream f(path);
    if (!f.is_open()) throw Error(CCP_ERR_IO, std::string("unable to open ") + path);
    f << "         avg   sqrt(var)         sem   f_n_b     i_t ter\n";
    for (uint32_t k = 0; k < 4; ++k) {
      const SeriesStat &t = a->sv[k].back();
      f << stat3(t) << cnfs(t.blocks, 8, ' ') << cnfs(t.cut, 8, ' ') << (t.settled ? " yes" : "  no") << "\n";
    }
    f << "\n\n";
    for (uint32_t ty = 0; ty < 3; ++ty) {
      f << "         r_b         avg   sqrt(var)         sem\n";
      f << "    0.000000    0.000000    0.000000    0.000000\n";
      for (uint32_t b = 0; b < nb; ++b)
        f << cnfs(h->R_n * std::pow((b + 1.0) / nb, 1.0 / 3), 12, ' ', 6) << stat3(a->sv[4 + ty * nb + b].back())
          << "\n";
      f << "\n\n";
    }
    for (auto *v : {&sd, &cp})
      for (uint32_t c = 0; c < 6; ++c) {
        f << "     s         avg\n";
        for (uint32_t i = 0; i < a->lsdcp[c]; ++i)
          f << cnfs(i + 1, 6, ' ') << cnfs((*v)[c][i] / a->n_dp, 12, ' ', 6) << "\n";
        f << "\n\n";
      }
    if (f.fail()) throw Error(CCP_ERR_IO, std::string("failed to write ") + path);
    // close the basic observables and clear the series (clear_is_var, chrana.cu:250-283)
    for (uint32_t c = 0; c < 6; ++c) {
      const uint32_t L = a->lsdcp[c];
      double *b = a->sdcp[c].p;
      launch_close_sums(h->stream, b, b + 2 * L, b + 4 * L, L, a->n_dp);
      launch_close_sums(h->stream, b + L, b + 3 * L, b + 5 * L, L, a->n_dp);
    }
    launch_close_sums(h->stream, a->cm.p, a->cm.p + a->lcm, a->cm.p + 2 * (size_t)a->lcm, a->lcm, a->n_dp);
    hip_check(hipGetLastError());
    hip_check(hipStreamSynchronize(h->stream));
    for (auto &t : a->ts) t.clear();
    a->n_dp = 0;
    ++a->n_s;
  });
}

int ccp_analysis_finish(ccp_analysis *a, const char *fin_path, const char *cm_path) {
  return aguarded(a, [&] {
    need(fin_path != nullptr && cm_path != nullptr, "null path");
    need(a->n_s > 0, "no simulations ended");
    ccp_sim *h = a->h;
    constexpr uint32_t nb = ccp_analysis::N_B;
    // calc_cs_final_stat + save_cs_final_stat (chrana.cu:286-360)
    std::ofstream f(fin_path);
    if (!f.is_open()) throw Error(CCP_ERR_IO, std::string("unable to open ") + fin_path);
    f << "         avg   sqrt(var)         sem\n";
    for (uint32_t k = 0; k < 4; ++k) f << stat3(weighted_stats(a->sv[k])) << "\n";
    f << "\n\n";
    for (uint32_t ty = 0; ty < 3; ++ty) {
      f << "         r_b         avg   sqrt(var)         sem\n";
      f << "    0.000000    0.000000    0.000000    0.000000\n";
      for (uint32_t b = 0; b < nb; ++b)
        f << cnfs(h->R_n * std::pow((b + 1.0) / nb, 1.0 / 3), 12, ' ', 6)
          << stat3(weighted_stats(a->sv[4 + ty * nb + b])) << "\n";
      f << "\n\n";
    }
    for (int which = 0; which < 2; ++which)
      for (uint32_t c = 0; c < 6; ++c) {
        const uint32_t L = a->lsdcp[c];
        std::vector<double> m1(L), m2(L);
        if (L) {
          hip_check(hipMemcpy(m1.data(), a->sdcp[c].p + (2 + which) * (size_t)L, L * 8, hipMemcpyDeviceToHost));
          hip_check(hipMemcpy(m2.data(), a->sdcp[c].p + (4 + which) * (size_t)L, L * 8, hipMemcpyDeviceToHost));
        }
        f << "     s         avg   sqrt(var)         sem\n";
        for (uint32_t i = 0; i < L; ++i) f << cnfs(i + 1, 6, ' ') << stat3(plain_stats(m1[i], m2[i], a->n_s)) << "\n";
        f << "\n\n";
      }
    if (f.fail()) throw Error(CCP_ERR_IO, std::string("failed to write ") + fin_path);
    // save_cm_avg_bin (chrana.cu:363-369): the average of each pixel
    std::vector<double> m1(a->lcm);
    hip_check(hipMemcpy(m1.data(), a->cm.p + a->lcm, (size_t)a->lcm * 8, hipMemcpyDeviceToHost));
    for (auto &v : m1) v = v / a->n_s;
    std::ofstream b(cm_path, std::ios::binary);
    if (!b.is_open()) throw Error(CCP_ERR_IO, std::string("unable to open ") + cm_path);
    b.write(reinterpret_cast<const char *>(m1.data()), (std::streamsize)m1.size() * 8);
    if (b.fail()) throw Error(CCP_ERR_IO, std::string("failed to write ") + cm_path);
  });
}

int ccp_sequence_pattern(const char *param_text, uint8_t *flags, size_t n) {
  try {
    ccp_sim one;
    one.log_fn = [](const char *, void *) {};
    ParMap par(param_text);
    one.configure(par);
    if (!flags || n != one.N) throw Error(CCP_ERR_ARG, "pattern size mismatch");
    // the draws of generate_initial_condition up to the bead types
    HostRng g(one.seed);
    one.set_lbs_positions(g);
    one.set_particle_types(g);
    std::memcpy(flags, one.hfl.data(), n);
    return CCP_OK;
  } catch (const Error &e) {
    g_create_error = e.what();
    return e.code;
  } catch (const std::exception &e) {
    g_create_error = e.what();
    return CCP_ERR_ARG;
  }
}

int ccp_series_stats(const float *v, size_t n, double out[6]) {
  if (!v || !out || n == 0) return CCP_ERR_ARG;
  const SeriesStat s = series_stats(v, n);
  out[0] = s.avg;
  out[1] = s.var;
  out[2] = s.sem;
  out[3] = s.blocks;
  out[4] = s.cut;
  out[5] = s.settled ? 1.0 : 0.0;
  return CCP_OK;
}

int ccp_weighted_stats(const double *avg, const double *sem, size_t n, double out[3]) {
  if (!avg || !sem || !out || n == 0) return CCP_ERR_ARG;
  std::vector<SeriesStat> v(n);
  for (size_t i = 0; i < n; ++i) {
    v[i].avg = avg[i];
    v[i].sem = sem[i];
  }
  const SeriesStat s = weighted_stats(v);
  out[0] = s.avg;
  out[1] = s.var;
  out[2] = s.sem;
  return CCP_OK;
}

} // extern "C"
