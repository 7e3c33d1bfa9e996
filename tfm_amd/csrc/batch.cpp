// Batch engine of libccpmi: many independent systems (one chain or a few
// each, their own nucleus size, eps_het, eps_lam, het/LAD pattern and seed)
// stepped together. Their beads share one index-ordered array and their
// cell grids one concatenated cell range, so one force launch and one order
// launch per Heun stage cover every system, and a hipGraph replays blocks of
// steps. Each system evolves exactly as it would alone with the same
// parameters (same cells, same pair order, same noise counters).
//
// Samples every contact_interval steps: a binned contact map per system
// (contact_bins^2 uint32, bins of ceil(n / contact_bins) beads) and its
// radial density (rd_bins shells of equal volume, all beads, ccp's lens
// formula and bead radius rco/2, chrana.cu:421-480).

#include "ccpmi.h"
#include "sim.hpp"

#include <cmath>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

namespace {

thread_local std::string g_batch_create_error;

// Host state of one uniform grid over the concatenated cell range.
struct BatchGrid {
  GridGeom g{}; // n_c, cap, nblk of the whole range
  DevBuf<uint32_t> count, bs0, bs1, cstart;
  DevBuf<float4> bucket, spos; // bucket: members' positions, index in the w bits
  DevBuf<uint32_t> hbs0, hbs1, work; // type-ordered work list (k_force_typed)
  int clean = 0;
  bool valid = false;
  uint32_t *bs(int k) { return k ? bs1.p : bs0.p; }
  uint32_t *hbs(int k) { return k ? hbs1.p : hbs0.p; }
  void alloc(size_t n_obj, bool typed) {
    count.alloc(g.n_c);
    bs0.alloc(g.nblk);
    bs1.alloc(g.nblk);
    bucket.alloc((size_t)g.n_c * g.cap);
    cstart.alloc((size_t)g.n_c + 1);
    spos.alloc(n_obj);
    hip_check(hipMemset(count.p, 0, (size_t)g.n_c * 4));
    hip_check(hipMemset(bs0.p, 0, (size_t)g.nblk * 4));
    hip_check(hipMemset(bs1.p, 0, (size_t)g.nblk * 4));
    hip_check(hipMemset(cstart.p, 0, ((size_t)g.n_c + 1) * 4));
    if (typed) {
      hbs0.alloc(g.nblk);
      hbs1.alloc(g.nblk);
      work.alloc(n_obj);
      hip_check(hipMemset(hbs0.p, 0, (size_t)g.nblk * 4));
      hip_check(hipMemset(hbs1.p, 0, (size_t)g.nblk * 4));
      hip_check(hipMemset(work.p, 0, std::max<size_t>(n_obj, 1) * 4));
    }
    clean = 0;
    valid = false;
  }
  void free_() {
    DevBuf<uint32_t> *u[] = {&count, &bs0, &bs1, &cstart, &hbs0, &hbs1, &work};
    for (auto *b : u) b->free_();
    bucket.free_();
    spos.free_();
  }
  OrderJob order_job(uint32_t n_het, uint32_t n_slots) {
    OrderJob j{g, count.p, bs(clean), bs(clean ^ 1), bucket.p, spos.p, nullptr, cstart.p};
    if (work.p) {
      j.work = work.p;
      j.hbs_read = hbs(clean);
      j.hbs_clear = hbs(clean ^ 1);
      j.n_het = n_het;
      j.n_slots = n_slots;
    }
    return j;
  }
};

} // namespace

struct ccp_batch {
  uint32_t n_sys = 0, N = 0, bins = 128, rd_bins = 32, cinterval = 0, graph_steps = 64;
  uint32_t spf = 0; // relaxation steps between overlap counts of the initial conditions
  bool integrate_pass = false;
  std::vector<std::string> text;     // parameter text of each system alone
  std::vector<SysParams> sys;
  std::vector<uint32_t> slot;        // n_sys + 1
  std::vector<uint32_t> nchain;      // chains per system
  KParams kp{};                      // shared constants; N = all beads
  float contact_rc = 1.5f;
  std::vector<float> hr;             // xyz by bead
  std::vector<uint8_t> hfl;          // CCP_HET / CCP_LAD by bead
  std::vector<uint32_t> hbond;       // W_B* bits by bead
  uint64_t step = 0;
  bool dstep_dirty = true;
  uint64_t noise_step = ~0ull; // step whose normals znoise holds
  uint32_t n_samples = 0;
  std::vector<double> rd_sum;        // n_sys x rd_bins
  // device
  int dev = 0;
  hipStream_t stream = nullptr;
  DevBuf<float4> r, er, pstar, fpair, znoise, fbond, rsave;
  DevBuf<SysParams> dsys;
  DevBuf<uint32_t> dslot, flags, cmaps;
  DevBuf<uint64_t> dstep;
  DevBuf<unsigned long long> dover; // overlap count per system
  DevBuf<double> obs;
  BatchGrid grid;
  hipGraphExec_t graphs[2] = {nullptr, nullptr};
  std::string err;

  void drop_graphs() {
    for (auto &g : graphs)
      if (g) {
        (void)hipGraphExecDestroy(g);
        g = nullptr;
      }
  }
  ~ccp_batch() {
    drop_graphs();
    DevBuf<float4> *f4[] = {&r, &er, &pstar, &fpair, &znoise, &fbond, &rsave};
    for (auto *b : f4) b->free_();
    DevBuf<uint32_t> *u4[] = {&dslot, &flags, &cmaps};
    for (auto *b : u4) b->free_();
    dsys.free_();
    dstep.free_();
    dover.free_();
    obs.free_();
    grid.free_();
    if (stream) (void)hipStreamDestroy(stream);
  }

  void upload() {
    uint32_t n_het = 0;
    for (uint32_t i = 0; i < N; ++i) n_het += (hfl[i] & CCP_HET) ? 1u : 0u;
    if (n_het != (uint32_t)kp.n_het) {
      // captured graphs hold the kernel parameters by value
      drop_graphs();
      set_het_count(kp, n_het);
    }
    std::vector<float4> h(N);
    for (uint32_t i = 0; i < N; ++i) {
      uint32_t w = i | hbond[i];
      if (hfl[i] & CCP_HET) w |= W_HET;
      if (hfl[i] & CCP_LAD) w |= W_LAD;
      float wf;
      std::memcpy(&wf, &w, 4);
      h[i] = make_float4(hr[3 * i], hr[3 * i + 1], hr[3 * i + 2], wf);
    }
    hip_check(hipMemcpyAsync(r.p, h.data(), (size_t)N * sizeof(float4), hipMemcpyHostToDevice, stream));
    hip_check(hipStreamSynchronize(stream));
    grid.valid = false;
  }
  void download() {
    std::vector<float4> h(N);
    hip_check(hipMemcpyAsync(h.data(), r.p, (size_t)N * sizeof(float4), hipMemcpyDeviceToHost, stream));
    hip_check(hipStreamSynchronize(stream));
    for (uint32_t i = 0; i < N; ++i) {
      hr[3 * i] = h[i].x;
      hr[3 * i + 1] = h[i].y;
      hr[3 * i + 2] = h[i].z;
    }
  }
  void check_flags() {
    uint32_t f[FLAG_COUNT];
    hip_check(hipMemcpyAsync(f, flags.p, sizeof(f), hipMemcpyDeviceToHost, stream));
    hip_check(hipStreamSynchronize(stream));
    if (f[FLAG_NONFINITE] || f[FLAG_OUTSIDE] || f[FLAG_OVERFLOW]) {
      hip_check(hipMemset(flags.p, 0, sizeof(f)));
      if (f[FLAG_NONFINITE]) throw Error(CCP_ERR_STATE, "non-finite bead position");
      if (f[FLAG_OUTSIDE]) throw Error(CCP_ERR_STATE, "bead left the cell grid");
      throw Error(CCP_ERR_STATE, "cell_capacity exceeded; raise cell_capacity");
    }
  }
  bool take_overflow() {
    uint32_t f = 0;
    hip_check(hipMemcpyAsync(&f, flags.p + FLAG_OVERFLOW, 4, hipMemcpyDeviceToHost, stream));
    hip_check(hipStreamSynchronize(stream));
    if (f) hip_check(hipMemset(flags.p + FLAG_OVERFLOW, 0, 4));
    return f != 0;
  }
  // cell list of the current positions, doubling the capacity until every
  // cell fits (outside the step loop)
  void ensure_grid() {
    if (grid.valid) return;
    for (;;) {
      launch_bin_batch(stream, dsys.p, dslot.p, (int)n_sys, (int)N, r.p, grid.count.p, grid.bs(grid.clean),
                       grid.bucket.p, flags.p, grid.work.p ? grid.hbs(grid.clean) : nullptr);
      const OrderJob j = grid.order_job((uint32_t)kp.n_het, N);
      launch_order(stream, j, nullptr, flags.p, nullptr);
      hip_check(hipGetLastError());
      grid.clean ^= 1;
      if (!take_overflow()) break;
      grow();
    }
    grid.valid = true;
  }
  // double the cells' storage (the grid is rebuilt on its next use)
  void grow() {
    if (grid.g.cap >= 4096) throw Error(CCP_ERR_STATE, "cell_capacity exceeded; raise cell_capacity");
    hip_check(hipStreamSynchronize(stream));
    drop_graphs();
    grid.g.cap *= 2;
    kp.gf.cap = grid.g.cap;
    for (auto &s : sys) s.g.cap = grid.g.cap;
    upload_sys();
    grid.bucket.free_();
    grid.bucket.alloc((size_t)grid.g.n_c * grid.g.cap);
    grid.valid = false;
  }
  void upload_sys() {
    hip_check(hipMemcpyAsync(dsys.p, sys.data(), sys.size() * sizeof(SysParams), hipMemcpyHostToDevice, stream));
    hip_check(hipStreamSynchronize(stream));
  }
  void sync_step() {
    if (dstep_dirty) {
      hip_check(hipMemcpyAsync(dstep.p, &step, sizeof(uint64_t), hipMemcpyHostToDevice, stream));
      hip_check(hipStreamSynchronize(stream));
      dstep_dirty = false;
    }
  }
  // normals of every system's beads, each under its own key and local index
  NoiseJob noise_job(uint32_t ahead) const {
    return NoiseJob{znoise.p, dstep.p, ahead, N, 0u, 0u, dsys.p, dslot.p, (int)n_sys};
  }
  bool bond_pass() const { return kp.lanes == 8 && !integrate_pass; }
  BondJob bond_job(const float4 *pos, int mode) const {
    return BondJob{fbond.p, pos, N, mode, dsys.p, dslot.p, (int)n_sys};
  }
  // the WCA/LJ set, or the noise-free soft-repulsion relaxation of the
  // initial conditions (as ccp_sim's mode_params)
  KParams mode_params(int mode) const {
    KParams k = kp;
    if (mode == MODE_ICG) k.sd = 0.0f;
    return k;
  }
  // before the first stage of a run: the current step's normals (unless held
  // already) and the bonded terms of the current positions
  void prime_terms(int mode = MODE_LJ) {
    sync_step();
    const NoiseJob nj = noise_job(0);
    const BondJob bj = bond_job(r.p, mode);
    const KParams k = mode_params(mode);
    launch_bead_terms(stream, noise_step != step ? &nj : nullptr, k, bond_pass() ? &bj : nullptr);
    noise_step = step;
  }
  void launch_stage(int stage, int mode = MODE_LJ) {
    const KParams k = mode_params(mode);
    KPtrs q{};
    q.fspos = grid.spos.p;
    q.fcstart = grid.cstart.p;
    q.pos = stage == 1 ? r.p : er.p;
    q.out0 = stage == 1 ? er.p : r.p;
    q.pstar = pstar.p;
    q.fcount = grid.count.p;
    q.fblocksum = grid.bs(grid.clean);
    q.fhblocksum = grid.work.p ? grid.hbs(grid.clean) : nullptr;
    q.work = grid.work.p;
    q.fbucket = grid.bucket.p;
    q.flags = flags.p;
    q.step = dstep.p;
    q.fpair = fpair.p;
    q.noise = znoise.p;
    q.fbond = bond_pass() ? fbond.p : nullptr;
    q.sys = dsys.p;
    q.sys_slot = dslot.p;
    q.n_sys = (int)n_sys;
    launch_force(stream, mode, stage, k, q);
    const OrderJob j = grid.order_job((uint32_t)kp.n_het, N);
    // with the order pass: the next step's normals after the first stage,
    // the bonded terms of q.out0 after both
    const NoiseJob nj = noise_job(1);
    const BondJob bj = bond_job(q.out0, mode);
    launch_order(stream, j, nullptr, flags.p, stage == 2 ? dstep.p : nullptr, stage == 1 ? &nj : nullptr, &k,
                 bond_pass() ? &bj : nullptr);
    grid.clean ^= 1;
  }
  void run_chunk(uint64_t n) {
    const uint64_t G = graph_steps;
    if (n >= G) {
      hipGraphExec_t &ge = graphs[grid.clean];
      const int c = grid.clean;
      if (!ge) {
        hipGraph_t g;
        hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        for (uint64_t s = 0; s < G; ++s) {
          launch_stage(1);
          launch_stage(2);
        }
        hip_check(hipStreamEndCapture(stream, &g));
        hip_check(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
        hip_check(hipGraphDestroy(g));
      }
      while (n >= G) {
        hip_check(hipGraphLaunch(ge, stream));
        n -= G;
      }
      grid.clean = c;
    }
    for (uint64_t s = 0; s < n; ++s) {
      launch_stage(1);
      launch_stage(2);
    }
    hip_check(hipGetLastError());
  }
  // every system's contacts and radial density, one launch each
  void sample() {
    hip_check(hipMemsetAsync(obs.p, 0, (size_t)n_sys * rd_bins * sizeof(double), stream));
    launch_contacts_batch(stream, dsys.p, dslot.p, (int)n_sys, (int)N, contact_rc, bins, grid.spos.p, grid.cstart.p,
                          cmaps.p);
    launch_radial_batch(stream, dsys.p, dslot.p, (int)n_sys, (int)N, (double)RCO / 2.0, (int)rd_bins, r.p, obs.p);
    hip_check(hipGetLastError());
    std::vector<double> a((size_t)n_sys * rd_bins);
    hip_check(hipMemcpyAsync(a.data(), obs.p, a.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    hip_check(hipStreamSynchronize(stream));
    for (uint32_t s = 0; s < n_sys; ++s) {
      const double R = sys[s].R_n, shell = 4.0 / 3.0 * M_PI * R * R * R / rd_bins;
      for (uint32_t k = 0; k < rd_bins; ++k) rd_sum[(size_t)s * rd_bins + k] += a[(size_t)s * rd_bins + k] / shell;
    }
    ++n_samples;
  }
  // Overlapping non-bonded pairs (closer than the initial-overlap distance)
  // of each listed system, on the batch grid.
  std::vector<unsigned long long> count_overlaps(const std::vector<uint32_t> &which) {
    ensure_grid();
    hip_check(hipMemsetAsync(dover.p, 0, n_sys * sizeof(unsigned long long), stream));
    for (uint32_t s : which) {
      const SysParams &S = sys[s];
      KParams k = kp;
      k.N = (int)S.n;
      k.gf = S.g;
      launch_overlaps(stream, k, grid.spos.p, grid.cstart.p, S.slot, MIS * MIS, dover.p + s);
    }
    hip_check(hipGetLastError());
    std::vector<unsigned long long> h(n_sys);
    hip_check(hipMemcpyAsync(h.data(), dover.p, n_sys * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
    hip_check(hipStreamSynchronize(stream));
    return h;
  }
  void set_frozen(const std::vector<uint32_t> &active) {
    for (auto &S : sys) S.frozen = 1;
    for (uint32_t s : active) sys[s].frozen = 0;
    upload_sys();
  }
  // chrsim::generate_initial_condition's relaxation (chrsim.cu:470-522) for
  // every system at once: rounds of spf noise-free soft-repulsion steps until
  // a system has no overlapping pair. A finished system is frozen (zero force:
  // its beads keep their positions) while the others go on. A round in which
  // a cell overflowed is undone and redone with larger cells, as ccp_sim's
  // relaxation does, so each system ends exactly as it would alone.
  void relax() {
    std::vector<uint32_t> active;
    for (uint32_t s = 0; s < n_sys; ++s) active.push_back(s);
    auto still = [&](const std::vector<unsigned long long> &po) {
      std::vector<uint32_t> next;
      for (uint32_t s : active)
        if (po[s] > 0) next.push_back(s);
      return next;
    };
    active = still(count_overlaps(active));
    set_frozen(active);
    uint32_t rounds = 0;
    while (!active.empty()) {
      if (++rounds > 10000) throw Error(CCP_ERR_STATE, "unable to remove bead overlaps");
      hip_check(hipMemcpyAsync(rsave.p, r.p, (size_t)N * sizeof(float4), hipMemcpyDeviceToDevice, stream));
      for (;;) {
        ensure_grid();
        prime_terms(MODE_ICG);
        for (uint32_t k = 0; k < spf; ++k) {
          launch_stage(1, MODE_ICG);
          launch_stage(2, MODE_ICG);
        }
        hip_check(hipGetLastError());
        step += spf;
        noise_step = step;
        if (!take_overflow()) break;
        hip_check(hipMemcpyAsync(r.p, rsave.p, (size_t)N * sizeof(float4), hipMemcpyDeviceToDevice, stream));
        step -= spf;
        dstep_dirty = true;
        grow();
      }
      check_flags();
      const std::vector<uint32_t> next = still(count_overlaps(active));
      if (next.size() != active.size()) {
        active = next;
        set_frozen(active);
      }
    }
    for (auto &S : sys) S.frozen = 0;
    upload_sys();
  }
  void run_steps(uint64_t n) {
    ensure_grid();
    prime_terms();
    while (n > 0) {
      uint64_t chunk = n;
      if (cinterval) chunk = std::min<uint64_t>(n, cinterval - step % cinterval);
      run_chunk(chunk);
      step += chunk;
      noise_step = step;
      n -= chunk;
      if (cinterval && step % cinterval == 0) sample();
    }
  }
};

namespace {

template <class F> int bguarded(ccp_batch *b, F &&fn) {
  if (!b) return CCP_ERR_ARG;
  try {
    hip_check(hipSetDevice(b->dev));
    fn();
    return CCP_OK;
  } catch (const Error &e) {
    b->err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    b->err = e.what();
    return CCP_ERR_ARG;
  }
}

void quiet(const char *, void *) {}

void bneed(bool ok, const char *what) {
  if (!ok) throw Error(CCP_ERR_ARG, what);
}

// Parameter text of system s alone: the common text first (its keys win,
// as in ccp's parameter map), then the system's own keys.
std::string system_text(const std::string &common, const ccp_batch_system &s, uint32_t bins) {
  const double rco = (double)RCO;
  const double R_n = rco / 2.0 + rco / 2.0 * std::cbrt((double)s.n_beads / (double)s.volume_fraction);
  char buf[1024];
  std::snprintf(buf, sizeof(buf),
                "number_of_particles %u\nnucleus_radius %.6f\nnumber_of_chains %u\nforce_field lj\n"
                "epsilon_het %.9g\nepsilon_lamina %.9g\nhet_fraction %.9g\nlad_fraction %.9g\nseed %llu\n"
                "number_of_lbs 0\ncontact_bins %u\ncontact_bin_size %u\nreplica %u\n",
                s.n_beads, R_n, s.n_chains ? s.n_chains : 1u, (double)s.eps_het, (double)s.eps_lam,
                (double)s.het_fraction, (double)s.lad_fraction, (unsigned long long)s.seed, bins,
                (s.n_beads + bins - 1) / bins, s.replica);
  return common + "\n" + buf;
}

// The half-cutoff cells at eight lanes per bead (reach 2, cells below
// 1.3 sigma) have run clean in batches of up to this many beads (256 chains of
// 2048). A command that ran a 1.1M-bead ensemble of this shape for 1e4 steps,
// after initial conditions generated on 16 host threads, faulted the GPU; its
// cause is not known, so larger batches of this shape are refused.
constexpr uint64_t HALF_CELL_BATCH_LIMIT = 524288;

// Everything of a batch that needs no device: each system's parameter text,
// constants, geometry and place in the shared bead and cell ranges.
void plan_batch(ccp_batch &bt, const char *common_params, const ccp_batch_system *systems, uint32_t n_sys) {
  ccp_batch *b = &bt;
  if (!systems || n_sys == 0) throw Error(CCP_ERR_ARG, "no systems");
  const std::string common = common_params ? common_params : "";
  ParMap cpar(common.c_str());
  b->bins = cpar.get<uint32_t>("contact_bins", 128u);
  b->rd_bins = cpar.get<uint32_t>("radial_bins", 32u);
  if (!(1 <= b->bins && b->bins <= 4096)) throw Error(CCP_ERR_PARAM, "contact_bins out of range");
  if (!(1 <= b->rd_bins && b->rd_bins <= 1024)) throw Error(CCP_ERR_PARAM, "radial_bins out of range");
  b->n_sys = n_sys;
  b->slot.push_back(0);
  int coff = 0;
  uint64_t total = 0;
  for (uint32_t s = 0; s < n_sys; ++s) {
    const ccp_batch_system &in = systems[s];
    if (!(in.n_beads >= 4 && in.volume_fraction > 0.0f && in.volume_fraction < 0.74f))
      throw Error(CCP_ERR_PARAM, "batch system out of range");
    b->text.push_back(system_text(common, in, b->bins));
    // the system's own configuration gives its geometry and constants
    ccp_sim one;
    one.log_fn = quiet;
    ParMap par(b->text.back().c_str());
    one.configure(par);
    one.fill_kparams();
    if (one.split) throw Error(CCP_ERR_PARAM, "fine_cell_size is not supported in batches");
    SysParams S{};
    S.R_n = one.R_n; S.R_o = one.R_o; S.R_b = one.R_b;
    S.noc = one.noc; S.boc = one.boc; S.nod = one.nod; S.bod = one.bod;
    S.eps_het = one.eps_het; S.eps_lam = one.eps_lam;
    S.seed_lo = one.kp.seed_lo;
    S.seed_hi = one.kp.seed_hi;
    S.slot = (uint32_t)total;
    S.n = one.N;
    S.g = one.gf.g;
    S.g.coff = coff;
    if (s == 0) {
      b->kp = one.kp;
      b->contact_rc = one.contact_rc;
      b->cinterval = one.cinterval;
      b->graph_steps = one.graph_steps;
      b->spf = one.spf;
      b->integrate_pass = one.integrate_pass;
    } else if (one.kp.lanes != b->kp.lanes || one.kp.reach != b->kp.reach || one.gf.g.csl != b->kp.gf.csl ||
               one.gf.g.cap != b->kp.gf.cap || one.kp.dt != b->kp.dt || one.kp.sd != b->kp.sd) {
      throw Error(CCP_ERR_PARAM, "batch systems must share the kernel shape, timestep and temperature");
    }
    one.hfl.assign(one.N, 0);
    for (uint32_t i = 0; i < one.N; ++i) b->hbond.push_back(one.wbits(i) & (W_B0 | W_B1 | W_B2 | W_B3));
    b->sys.push_back(S);
    b->nchain.push_back((uint32_t)one.chain_off.size() - 1);
    // cell indices stay below 2^30 so that their byte offsets (index x 4)
    // fit the 32-bit offsets of the force kernels' row loads
    if ((uint64_t)coff + (uint64_t)one.gf.g.n_c >= (1ull << 30)) throw Error(CCP_ERR_PARAM, "batch too large");
    coff += one.gf.g.n_c;
    total += one.N;
    if (total >= (1ull << ID_BITS)) throw Error(CCP_ERR_PARAM, "batch too large");
    b->slot.push_back((uint32_t)total);
  }
  if (b->kp.lanes == 8 && b->kp.reach == 2 && b->kp.gf.csl < 1.3f && total > HALF_CELL_BATCH_LIMIT)
    throw Error(CCP_ERR_PARAM, "lanes_per_bead 8 with cell_reach 2 and cells below 1.3 sigma is limited to "
                               "524288 beads per batch: a larger run of this shape faulted the GPU and the cause "
                               "is not yet found");
  b->N = (uint32_t)total;
  b->kp.N = (int)b->N;
  b->kp.n_l = 0;
  b->kp.split = 0;
  b->kp.gh.n_c = 0;
  b->grid.g = b->kp.gf;
  b->grid.g.n_c = coff;
  b->grid.g.coff = 0;
  b->grid.g.nblk = (coff + SCAN_BLOCK - 1) / SCAN_BLOCK;
  b->hr.assign(3 * (size_t)b->N, 0.0f);
  b->hfl.assign(b->N, 0);
  b->rd_sum.assign((size_t)n_sys * b->rd_bins, 0.0);
}

} // namespace

extern "C" {

ccp_batch *ccp_batch_create(const char *common_params, const ccp_batch_system *systems, uint32_t n_sys) {
  try {
    std::unique_ptr<ccp_batch> b(new ccp_batch());
    plan_batch(*b, common_params, systems, n_sys);
    // device
    hip_check(hipGetDevice(&b->dev));
    hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    b->r.alloc(b->N);
    b->er.alloc(b->N);
    b->pstar.alloc(b->N);
    b->rsave.alloc(b->N);
    b->znoise.alloc(b->N);
    if (b->bond_pass()) b->fbond.alloc(b->N);
    if (b->kp.lanes == 8 && b->integrate_pass) b->fpair.alloc(b->N);
    b->dsys.alloc(n_sys);
    b->dslot.alloc(n_sys + 1);
    b->flags.alloc(FLAG_COUNT);
    b->cmaps.alloc((size_t)n_sys * b->bins * b->bins);
    b->dstep.alloc(1);
    b->dover.alloc(n_sys);
    b->obs.alloc((6 + 3 * (size_t)b->rd_bins) * n_sys);
    b->grid.alloc(b->N, b->kp.typed != 0);
    hip_check(hipMemcpy(b->dsys.p, b->sys.data(), n_sys * sizeof(SysParams), hipMemcpyHostToDevice));
    hip_check(hipMemcpy(b->dslot.p, b->slot.data(), (n_sys + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
    hip_check(hipMemset(b->flags.p, 0, FLAG_COUNT * sizeof(uint32_t)));
    hip_check(hipMemset(b->cmaps.p, 0, (size_t)n_sys * b->bins * b->bins * sizeof(uint32_t)));
    hip_check(hipMemset(b->dstep.p, 0, sizeof(uint64_t)));
    b->upload();
    return b.release();
  } catch (const std::exception &e) {
    g_batch_create_error = e.what();
    return nullptr;
  }
}

int ccp_batch_layout(const char *common_params, const ccp_batch_system *systems, uint32_t n_sys,
                     uint32_t *n_beads, uint64_t *n_cells, uint64_t *bucket_bytes, uint64_t *cell_offsets) {
  try {
    ccp_batch b;
    plan_batch(b, common_params, systems, n_sys);
    if (n_beads) *n_beads = b.N;
    if (n_cells) *n_cells = (uint64_t)b.grid.g.n_c;
    if (bucket_bytes) *bucket_bytes = (uint64_t)b.grid.g.n_c * (uint64_t)b.grid.g.cap * sizeof(float4);
    if (cell_offsets) {
      for (uint32_t s = 0; s < n_sys; ++s) cell_offsets[s] = (uint64_t)b.sys[s].g.coff;
      cell_offsets[n_sys] = (uint64_t)b.grid.g.n_c;
    }
    return CCP_OK;
  } catch (const Error &e) {
    g_batch_create_error = e.what();
    return e.code;
  } catch (const std::exception &e) {
    g_batch_create_error = e.what();
    return CCP_ERR_ARG;
  }
}

void ccp_batch_destroy(ccp_batch *b) {
  if (b) (void)hipSetDevice(b->dev);
  delete b;
}

const char *ccp_batch_last_error(const ccp_batch *b) { return b ? b->err.c_str() : g_batch_create_error.c_str(); }

int ccp_batch_info(const ccp_batch *b, uint32_t *n_sys, uint32_t *n_beads, uint32_t *n_cells, uint32_t *offsets) {
  if (!b) return CCP_ERR_ARG;
  if (n_sys) *n_sys = b->n_sys;
  if (n_beads) *n_beads = b->N;
  if (n_cells) *n_cells = (uint32_t)b->grid.g.n_c;
  if (offsets)
    for (uint32_t s = 0; s <= b->n_sys; ++s) offsets[s] = b->slot[s];
  return CCP_OK;
}

int ccp_batch_system_params(const ccp_batch *b, uint32_t s, char *buf, size_t len) {
  if (!b || s >= b->n_sys || !buf || len == 0) return CCP_ERR_ARG;
  const std::string &t = b->text[s];
  if (t.size() + 1 > len) return CCP_ERR_ARG;
  std::memcpy(buf, t.c_str(), t.size() + 1);
  return CCP_OK;
}

// Each system's initial condition is the one a ccp_sim of its parameters
// generates alone: its host walk, then the relaxation, run here for all
// systems together on the batch grid.
int ccp_batch_generate_initial_conditions(ccp_batch *b) {
  return bguarded(b, [&] {
    for (uint32_t s = 0; s < b->n_sys; ++s) {
      ccp_sim one;
      one.log_fn = quiet;
      ParMap par(b->text[s].c_str());
      one.configure(par);
      HostRng g(one.seed);
      one.set_lbs_positions(g);
      one.set_particle_types(g);
      if (one.replica) {
        HostRng w(one.seed, one.replica);
        one.set_particle_positions(w);
      } else {
        one.set_particle_positions(g);
      }
      const uint32_t o = b->slot[s];
      std::copy(one.hr.begin(), one.hr.end(), b->hr.begin() + 3 * (size_t)o);
      std::copy(one.hfl.begin(), one.hfl.end(), b->hfl.begin() + o);
    }
    hip_check(hipSetDevice(b->dev));
    b->upload();
    b->step = 0;
    b->dstep_dirty = true;
    b->relax();
    b->download();
    b->step = 0;
    b->dstep_dirty = true;
  });
}

int ccp_batch_run_steps(ccp_batch *b, uint64_t n) {
  return bguarded(b, [&] { b->run_steps(n); });
}

int ccp_batch_synchronize(ccp_batch *b) {
  return bguarded(b, [&] {
    hip_check(hipStreamSynchronize(b->stream));
    b->check_flags();
  });
}

int ccp_batch_get_state(ccp_batch *b, float *xyz, uint8_t *flags, size_t n) {
  return bguarded(b, [&] {
    bneed(n == b->N, "state size mismatch");
    b->download();
    b->check_flags();
    if (xyz) std::memcpy(xyz, b->hr.data(), 12 * (size_t)b->N);
    if (flags) std::memcpy(flags, b->hfl.data(), b->N);
  });
}

int ccp_batch_set_state(ccp_batch *b, const float *xyz, const uint8_t *flags, size_t n) {
  return bguarded(b, [&] {
    bneed(n == b->N, "state size mismatch");
    if (xyz) std::memcpy(b->hr.data(), xyz, 12 * (size_t)b->N);
    if (flags) std::memcpy(b->hfl.data(), flags, b->N);
    b->upload();
  });
}

int ccp_batch_get_step(const ccp_batch *b, uint64_t *step) {
  if (!b || !step) return CCP_ERR_ARG;
  *step = b->step;
  return CCP_OK;
}

int ccp_batch_set_step(ccp_batch *b, uint64_t step) {
  if (!b) return CCP_ERR_ARG;
  b->step = step;
  b->dstep_dirty = true;
  return CCP_OK;
}

int ccp_batch_contact_maps(ccp_batch *b, uint32_t *maps, size_t n, uint32_t *n_samples) {
  return bguarded(b, [&] {
    bneed(maps && n == (size_t)b->n_sys * b->bins * b->bins, "contact map size mismatch");
    hip_check(hipMemcpyAsync(maps, b->cmaps.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, b->stream));
    hip_check(hipStreamSynchronize(b->stream));
    if (n_samples) *n_samples = b->n_samples;
  });
}

int ccp_batch_radial_density(ccp_batch *b, double *rd, size_t n) {
  return bguarded(b, [&] {
    bneed(rd && n == (size_t)b->n_sys * b->rd_bins, "radial density size mismatch");
    for (size_t k = 0; k < n; ++k) rd[k] = b->n_samples ? b->rd_sum[k] / b->n_samples : 0.0;
  });
}

int ccp_batch_reset_samples(ccp_batch *b) {
  return bguarded(b, [&] {
    hip_check(hipMemsetAsync(b->cmaps.p, 0, (size_t)b->n_sys * b->bins * b->bins * sizeof(uint32_t), b->stream));
    hip_check(hipStreamSynchronize(b->stream));
    std::fill(b->rd_sum.begin(), b->rd_sum.end(), 0.0);
    b->n_samples = 0;
  });
}

int ccp_batch_time_steps(ccp_batch *b, uint32_t n, float *ms_per_step) {
  return bguarded(b, [&] {
    bneed(ms_per_step && n > 0, "bad arguments");
    b->ensure_grid();
    b->prime_terms();
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0));
    hip_check(hipEventCreate(&e1));
    hip_check(hipEventRecord(e0, b->stream));
    for (uint32_t s = 0; s < n; ++s) {
      b->launch_stage(1);
      b->launch_stage(2);
    }
    hip_check(hipEventRecord(e1, b->stream));
    hip_check(hipEventSynchronize(e1));
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, e0, e1));
    *ms_per_step = ms / n;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->step += n;
    b->noise_step = b->step;
    b->check_flags();
  });
}

} // extern "C"
