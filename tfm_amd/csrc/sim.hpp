// Internal definition of the simulation handle (struct ccp_sim) and the
// host helpers of its initial-condition generator, shared by the C ABI
// (ccpmi.cpp) and the batch engine (batch.cpp).
#pragma once

#include "host.hpp"
#include "kernels.hpp"

#include <cmath>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <map>
#include <string>
#include <vector>

namespace ccpmi {
namespace host {


// ---------------------------------------------------------------- host Philox stream
// Uniforms for the initial condition: Philox4x32-10 with the simulation seed
// as key and counter {n lo, n hi, 0xffffffff, 1 << 16}, disjoint from the
// step-noise counters {bead, step lo, step hi, attempt < 2^16}.
struct HostRng {
  uint32_t k0, k1;
  uint64_t n = 0;
  uint32_t buf[4];
  int pos = 4;
  uint32_t stream = 1u << 16; // fourth counter word; replicas > 0 use their own stream
  explicit HostRng(uint64_t seed, uint32_t replica = 0)
      : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)), stream((1u << 16) + replica) {}
  void refill() {
    uint32_t c[4] = {(uint32_t)n, (uint32_t)(n >> 32), 0xffffffffu, stream};
    uint32_t a = k0, b = k1;
    for (int r = 0; r < 10; ++r) {
      uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
      uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ a, n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ b;
      c[1] = (uint32_t)p1;
      c[3] = (uint32_t)p0;
      c[0] = n0;
      c[2] = n2;
      a += 0x9E3779B9u;
      b += 0xBB67AE85u;
    }
    std::memcpy(buf, c, sizeof(c));
    ++n;
    pos = 0;
  }
  // uniform in (0, 1] like curandGenerateUniform
  float uniform() {
    if (pos == 4) refill();
    return (float)((buf[pos++] >> 8) + 1u) * (1.0f / 16777216.0f);
  }
};

// inverse error function: Winitzki's closed-form start (2008) refined by
// Newton steps on std::erf
inline double erfinv(double x) {
  if (x <= -1.0) return -INFINITY;
  if (x >= 1.0) return INFINITY;
  const double a = 0.147, ln = std::log(1.0 - x * x), t = 2.0 / (M_PI * a) + ln / 2.0;
  double y = std::copysign(std::sqrt(std::sqrt(t * t - ln / a) - t), x);
  for (int k = 0; k < 4; ++k) y -= (std::erf(y) - x) / (2.0 / std::sqrt(M_PI) * std::exp(-y * y));
  return y;
}

struct V3 {
  float x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }


} // namespace host
} // namespace ccpmi

// internal header: the library's own sources use these names unqualified
using namespace ccpmi;
using namespace ccpmi::host;


// ---------------------------------------------------------------- handle
struct ccp_sim {
  // parameters
  uint32_t N = 0, n_l = 0, fpf = 0, spf = 0, tpb = 0;
  float T = 0, R_n = 0, R_o = 0, R_b = 0;
  float noc = 0, boc = 0, nod = 0, bod = 0, d_m = 0;
  // pair set (MODE_*) whose bonded and wall terms fbond holds for the
  // positions in r (-1: none, or stale after a host-side change of r)
  int terms_mode = -1;
  float dt = 0, sd = 0;
  uint64_t seed = 0;
  uint32_t replica = 0; // > 0: same bead types as replica 0, its own walk and noise
  int ff = CCP_FF_CCP;
  float csl = 0;   // coarse cell side: het grid and ccp-layout export grid
  float csl_f = 0; // cell side of the all-bead grid of the step (fine when split)
  bool split = false; // short-range pairs on the fine grid + het-het tail on the het grid
  std::vector<uint32_t> chain_off;
  std::string seq_file;
  int seq_mode = 0; // 0 ccp, 1 het/LAD generator
  float het_fraction = 0.3f, het_mean = 20.0f, lad_fraction = 0.1f;
  int lad_ends = 1;
  float k_b = K_BEND;
  int bend_all = 0, bond_type = 0;
  float eps_het = 0.5f, eps_lam = 1.0f;
  float contact_rc = 0;
  uint32_t cbins = 0, cbin_size = 0, cinterval = 0;
  uint32_t graph_steps = 64;
  bool use_graphs = true; // false: every step launched from the host (profilers that cannot follow graphs)
  uint32_t reach = 1; // neighbour cells scanned on each side
  uint32_t lanes = 3; // lanes per bead of the force kernel
  bool integrate_pass = false; // 8 lanes: bonds, wall and stage update in a one-lane-per-bead pass
  bool group_types = true;     // 8 lanes: het beads first within each workgroup
  bool typed = false;          // 8 lanes at reach 2: type-ordered work list (k_force_typed)
  uint32_t rounds = 1;         // 8 lanes: pair-phase rounds per wavefront before the per-bead work
  float fcap = 0.0f;
  KParams kp{};

  // state
  uint32_t i_f = 0;
  float t = 0.0f;
  uint64_t step = 0;
  std::vector<float> hr, hlr;
  std::vector<uint8_t> hfl;
  bool lgrid_valid = false, dstep_dirty = true;
  uint32_t n_samples = 0;
  // step whose normals znoise holds (the first Heun stage reads them; the
  // order pass after it writes the next step's)
  uint64_t noise_step = ~0ull;

  // device
  int dev = 0;
  hipStream_t stream = nullptr;
  DevBuf<float4> r, er, pstar, lr, lspos, aux, fpair, znoise, fbond;
  DevBuf<uint32_t> lstart, flags, cmap;
  // completion word and flags written by k_signal: coherent pinned host
  // memory (host and device addresses) and the last sequence number sent
  uint32_t *hsig = nullptr, *hsig_dev = nullptr;
  uint32_t sig_seq = 0;
  DevBuf<uint64_t> dstep;
  DevBuf<unsigned long long> dcount;
  // one uniform grid: binning buffers + ordered cell list
  struct HostGrid {
    GridGeom g{};
    DevBuf<uint32_t> count, bs0, bs1, cstart, soi;
    DevBuf<float4> bucket, spos; // bucket: members' positions, index in the w bits
    // type-ordered work list and the per-block het totals of the binning
    // (double buffered as bs0/bs1), for the step grid of k_force_typed
    DevBuf<uint32_t> hbs0, hbs1, work;
    int clean = 0;       // which per-block total buffer is zero
    bool valid = false;  // holds the cell list of the current positions
    bool allocated = false;
    uint32_t *bs(int k) { return k ? bs1.p : bs0.p; }
    uint32_t *hbs(int k) { return k ? hbs1.p : hbs0.p; }
    void alloc(size_t n_obj, bool typed = false) {
      count.alloc(g.n_c); bs0.alloc(g.nblk); bs1.alloc(g.nblk);
      bucket.alloc((size_t)g.n_c * g.cap); cstart.alloc(g.n_c + 1); soi.alloc(n_obj); spos.alloc(n_obj);
      hip_check(hipMemset(count.p, 0, (size_t)g.n_c * 4));
      hip_check(hipMemset(bs0.p, 0, (size_t)g.nblk * 4));
      hip_check(hipMemset(bs1.p, 0, (size_t)g.nblk * 4));
      hip_check(hipMemset(cstart.p, 0, ((size_t)g.n_c + 1) * 4));
      if (typed) {
        hbs0.alloc(g.nblk); hbs1.alloc(g.nblk); work.alloc(n_obj);
        hip_check(hipMemset(hbs0.p, 0, (size_t)g.nblk * 4));
        hip_check(hipMemset(hbs1.p, 0, (size_t)g.nblk * 4));
        // every entry a valid slot from the start
        hip_check(hipMemset(work.p, 0, std::max<size_t>(n_obj, 1) * 4));
      }
      clean = 0;
      valid = false;
      allocated = true;
    }
    void free_() {
      DevBuf<uint32_t> *u[] = {&count, &bs0, &bs1, &cstart, &soi, &hbs0, &hbs1, &work};
      for (auto *b : u) b->free_();
      bucket.free_();
      spos.free_();
      allocated = false;
    }
    // the ordering job of this grid's buffers; with the work list when the
    // grid has one and `typed` asks for it (n_het het beads of n_slots)
    OrderJob order_job(const GridGeom &geom, bool typed, uint32_t n_het, uint32_t n_slots, float4 *out_spos,
                       uint32_t *out_cstart, uint32_t *out_soi) {
      OrderJob j{geom, count.p, bs(clean), bs(clean ^ 1), bucket.p, out_spos, out_soi, out_cstart};
      if (work.p) {
        j.hbs_clear = hbs(clean ^ 1);
        if (typed) {
          j.work = work.p;
          j.hbs_read = hbs(clean);
          j.n_het = n_het;
          j.n_slots = n_slots;
        }
      }
      return j;
    }
  };
  HostGrid gf; // fine grid of every bead
  GridGeom lg{}; // LBS sites on the fine geometry (uses gf's binning buffers)
  HostGrid gh; // coarse grid of het beads
  HostGrid gx; // coarse grid of every bead in ccp's geometry (cell-list export)
  // Captured step blocks by (mode, clean buffer of each grid, length): the
  // length is graph_steps or a power of two below it, so a run of any length
  // replays graphs only.
  std::map<uint32_t, hipGraphExec_t> graphs;
  static uint32_t graph_key(int mode, int cf, int ch, uint32_t len) {
    return (uint32_t)mode | (uint32_t)cf << 2 | (uint32_t)ch << 3 | len << 4;
  }

  // misc
  std::string err;
  ccp_log_fn log_fn = nullptr;
  void *log_user = nullptr;

  void drop_graphs() {
    for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
  }

  ~ccp_sim() {
    drop_graphs();
    DevBuf<float4> *f4[] = {&r, &er, &pstar, &lr, &lspos, &aux, &fpair, &znoise, &fbond};
    for (auto *b : f4) b->free_();
    DevBuf<uint32_t> *u4[] = {&lstart, &flags, &cmap};
    for (auto *b : u4) b->free_();
    gf.free_();
    gh.free_();
    gx.free_();
    dstep.free_();
    dcount.free_();
    if (hsig) (void)hipHostFree(hsig);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // -------------------------------------------------------------- logging
  void record(const std::string &msg) {
    if (log_fn) {
      log_fn(msg.c_str(), log_user);
      return;
    }
    time_t now = time(nullptr);
    char ts[21];
    strftime(ts, 21, "[%d/%m/%y %H:%M:%S] ", localtime(&now));
    std::cout << ts << msg << std::endl;
  }
  void progress(float pc) {
    std::string line = "progress: " + cnfs(pc, 5, '0', 1) + "%";
    if (log_fn) {
      log_fn(line.c_str(), log_user);
      return;
    }
    std::cout << line << "\r";
    std::cout.flush();
  }

  // -------------------------------------------------------------- setup
  void configure(const ParMap &par);
  void allocate();
  void fill_kparams();
  uint32_t wbits(uint32_t i) const;
  void upload_positions();
  void upload_lbs();
  void download_positions();
  void check_flags();
  // -------------------------------------------------------------- step machinery
  KParams mode_params(int mode) const {
    KParams k = kp;
    if (mode == MODE_ICG) k.sd = 0.0f;
    return k;
  }
  void build_grid(HostGrid &G, const GridGeom &geom, const float4 *pos, int n_obj, bool het_only, float4 *out_spos,
                  uint32_t *out_cstart, uint32_t *out_soi, bool work_list = false);
  bool take_overflow();
  void grow_capacity();
  void ensure_grids();
  void invalidate() {
    gf.valid = gh.valid = gx.valid = false;
    terms_mode = -1;
  }
  void sync_step();
  NoiseJob noise_job(uint32_t ahead) const {
    return NoiseJob{znoise.p, dstep.p, ahead, N, kp.seed_lo, kp.seed_hi, nullptr, nullptr, 0};
  }
  // balanced kernel: bonds, bending and wall come from fbond, which the order
  // pass after each stage fills for the next one
  bool bond_pass() const { return lanes == 8 && !integrate_pass; }
  BondJob bond_job(const float4 *pos, int mode) const {
    return BondJob{fbond.p, pos, N, mode, nullptr, nullptr, 0};
  }
  // before the first stage of a run: the normals of the current step and the
  // bonded terms of the current positions, unless held already (the order
  // pass after each stage leaves both for the next one)
  void prime_terms(int mode) {
    sync_step();
    const NoiseJob nj = noise_job(0);
    const BondJob bj = bond_job(r.p, mode);
    const KParams k = mode_params(mode);
    const bool noise = noise_step != step, bond = bond_pass() && terms_mode != mode;
    launch_bead_terms(stream, noise ? &nj : nullptr, k, bond ? &bj : nullptr);
    noise_step = step;
    terms_mode = bond_pass() ? mode : -1;
  }
  // `mid` (optional) is recorded between the force and the order launches
  void launch_stage(int mode, int stage, hipEvent_t mid = nullptr);
  void launch_step(int mode) {
    launch_stage(mode, 1);
    launch_stage(mode, 2);
  }
  void run_steps(uint64_t n, int mode);
  void run_chunk(uint64_t n, int mode);
  hipGraphExec_t step_graph(uint32_t len, int mode);
  // every block length run_chunk can replay, captured ahead of a timed run
  void capture_graphs(int mode);
  // -------------------------------------------------------------- initial condition
  void set_lbs_positions(HostRng &g);
  void set_particle_types(HostRng &g);
  void set_particle_positions(HostRng &g);
  void perform_random_walk(HostRng &g, uint32_t i_s, uint32_t i_e, V3 dir, float mda);
  uint64_t count_overlaps(float dmin);
  void generate_initial_condition();
  // -------------------------------------------------------------- files
  void write_frame_bin(std::ofstream &f);
  // -------------------------------------------------------------- analysis
  // bead range of ccp's chromosome c for d(s)/P(s) (chrana.cu:93-97,484-491)
  bool sd_cp_range(uint32_t c, uint32_t &i_s, uint32_t &i_e) const {
    if (N == N_DEF) {
      if (c >= 6) return false;
      i_s = CHR_LIM[c];
      i_e = CHR_LIM[c + 1];
      return true;
    }
    if (c != 0) return false;
    i_s = 0;
    i_e = N;
    return true;
  }
};
