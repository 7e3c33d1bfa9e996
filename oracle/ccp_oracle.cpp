// CPU oracle for the ccp Brownian-dynamics step — TEST INFRASTRUCTURE ONLY.
// Scalar restatement of the reference algorithm; see ccp_oracle.h for scope,
// the reference lines it follows and how each function is pinned.
//
// Build: oracle/Makefile (g++ -O2 -fopenmp -ffp-contract=off).

#include "ccp_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------- vectors
template <class R> struct V3 {
  R x, y, z;
};
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline void operator+=(V3<R> &a, V3<R> b) { a.x += b.x; a.y += b.y; a.z += b.z; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline R len(V3<R> a) { return std::sqrt(dot(a, a)); }

// ---------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplies with the
// published multipliers and Weyl key increments.
constexpr uint32_t PH_M0 = 0xD2511F53u, PH_M1 = 0xCD9E8D57u;
constexpr uint32_t PH_W0 = 0x9E3779B9u, PH_W1 = 0xBB67AE85u;

inline void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int round = 0; round < 10; ++round) {
    uint64_t p0 = (uint64_t)PH_M0 * c[0];
    uint64_t p1 = (uint64_t)PH_M1 * c[2];
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
    k0 += PH_W0;
    k1 += PH_W1;
  }
}

// 24-bit uniform in (0, 1]: ((x >> 8) + 1) * 2^-24, exact in FP32.
inline double u24(uint32_t x) { return (double)((x >> 8) + 1u) * (1.0 / 16777216.0); }

// Three standard normals for (seed, bead, step) with ccp's rule of redrawing
// the whole triple while any |z| > 5 (chrsim.cu:350-356). Counter layout:
// {bead, step lo, step hi, attempt}; key {seed lo, seed hi}. Box–Muller on
// (u0,u1) -> z0,z1 and (u2,u3) -> z2.
int noise_triple(uint64_t seed, uint32_t bead, uint64_t step, double z[3]) {
  for (uint32_t a = 0;; ++a) {
    uint32_t c[4] = {bead, (uint32_t)step, (uint32_t)(step >> 32), a};
    philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    double r01 = std::sqrt(-2.0 * std::log(u24(c[0])));
    double t1 = 2.0 * M_PI * u24(c[1]);
    double r23 = std::sqrt(-2.0 * std::log(u24(c[2])));
    double t3 = 2.0 * M_PI * u24(c[3]);
    z[0] = r01 * std::cos(t1);
    z[1] = r01 * std::sin(t1);
    z[2] = r23 * std::cos(t3);
    if (!(std::fabs(z[0]) > 5 || std::fabs(z[1]) > 5 || std::fabs(z[2]) > 5)) return (int)a;
    if (a == 0xFFFFu) return (int)a;
  }
}

// ---------------------------------------------------------------- system
// Nucleus geometry (chrdat.cu:13-19): derived in double, stored in R.
template <class R> struct Geo {
  R R_n, R_o, R_b, noc, boc, nod, bod, d_b, d_m;
};
template <class R> Geo<R> make_geo(const orc_params *p) {
  double Rn = p->R_n, Ro = p->R_o, Rb = p->R_b;
  double noc = std::cos(std::asin(Ro / Rn));
  double boc = std::cos(std::asin(Ro / Rb)); // NaN when R_b = 0, as in ccp
  double nod = std::sqrt(Rn * Rn - Ro * Ro);
  double bod = std::sqrt(Rb * Rb - Ro * Ro);
  Geo<R> g;
  g.R_n = (R)Rn; g.R_o = (R)Ro; g.R_b = (R)Rb;
  g.noc = (R)noc; g.boc = (R)boc; g.nod = (R)nod; g.bod = (R)bod;
  g.d_b = (R)(nod + bod); g.d_m = (R)(nod + bod + Rb);
  return g;
}

template <class R> struct Sys {
  const orc_params *p;
  const uint8_t *fl;
  int N, n_l;
  std::vector<int32_t> lim_l, lim_u; // per bead bond limits (chrsim.cu:27-40)
  Geo<R> g;
  std::vector<V3<R>> lbs;
  // grid
  R csl;
  uint32_t cps, n_c;
};

template <class R>
Sys<R> make_sys(const orc_params *p, const uint8_t *fl, const int32_t *co, const double *lbs) {
  Sys<R> s;
  s.p = p;
  s.fl = fl;
  s.N = p->N;
  s.n_l = p->n_l;
  s.lim_l.resize(s.N);
  s.lim_u.resize(s.N);
  for (int c = 0; c < p->n_chains; ++c)
    for (int i = co[c]; i < co[c + 1]; ++i) {
      s.lim_l[i] = co[c] + 2;
      s.lim_u[i] = co[c + 1];
    }
  s.g = make_geo<R>(p);
  s.lbs.resize(s.n_l);
  for (int l = 0; l < s.n_l; ++l) s.lbs[l] = {(R)lbs[3 * l], (R)lbs[3 * l + 1], (R)lbs[3 * l + 2]};
  s.csl = (R)p->csl;
  s.cps = (uint32_t)p->cps;
  s.n_c = s.cps * s.cps * s.cps;
  return s;
}

// WCA / LJ scalar factor f such that force = f * (r_i - r_j):
// 24 eps s6 (2 s6 - 1) / d2 with s6 = (sigma^2/d2)^3.
template <class R> inline R lj_fac(R eps, R sig2, R d2) {
  R s2 = sig2 / d2;
  R s6 = s2 * s2 * s2;
  return (R)24 * eps * s6 * ((R)2 * s6 - (R)1) / d2;
}

// ccp's Wang-Frenkel factor e (18 d^4 - 96 d^2 + 96) / d^8 with its
// arithmetic (calc_ppf<DST>, chrsim.cu:184; calc_wpf, :87): d2 = d * d in R,
// the polynomial and the division in double (ccp's double literals), d^8 the
// R product ((d2 d2) d2) d2, the factor rounded to R. In FP64 every step is
// double, so this is the plain formula.
template <class R> inline R ccp_wf(R e, R d) {
  const R d2 = d * d;
  const double x = (double)d2;
  const double poly = 18.0 * x * x - 96.0 * x + 96.0;
  const R d8 = d2 * d2 * d2 * d2;
  return (R)((double)e * poly / (double)d8);
}

// ---------------------------------------------------------------- forces
// Bonded forces on bead i: harmonic (or FENE) bonds and cosine bending,
// following calc_bf (chrsim.cu:13-76) with per-chain limits.
template <class R> V3<R> bonded(const Sys<R> &S, const V3<R> *r, int i) {
  const orc_params *p = S.p;
  V3<R> vec[4];
  R il[4], cs[3];
  int lim_l = S.lim_l[i], lim_u = S.lim_u[i];
  for (int b = 0; b < 4; ++b) {
    int ib = i + b;
    if (ib >= lim_l && ib <= lim_u) {
      vec[b] = r[i + b - 1] - r[i + b - 2];
      il[b] = (R)1 / std::sqrt(dot(vec[b], vec[b]));
    } else {
      vec[b] = {0, 0, 0};
      il[b] = 0;
    }
  }
  for (int c = 0; c < 3; ++c) cs[c] = dot(vec[c + 1], vec[c]) * il[c + 1] * il[c];
  V3<R> bf = {0, 0, 0};
  if (p->bond_type == 0) {
    R ke = (R)p->k_e, l0 = (R)p->l_0;
    bf += ke * (((R)1 - l0 * il[2]) * vec[2] - ((R)1 - l0 * il[1]) * vec[1]);
  } else {
    R K = (R)p->fene_k, R02 = (R)(p->fene_r0 * p->fene_r0);
    R sig2 = (R)(p->sigma * p->sigma), rm2 = (R)(std::pow(2.0, 1.0 / 3.0) * p->sigma * p->sigma);
    R ew = (R)p->eps_wca;
    if (il[2] != 0) {
      R d2 = dot(vec[2], vec[2]);
      bf += (K / std::max((R)1 - d2 / R02, (R)0.01)) * vec[2];
      if (d2 < rm2) bf += (-lj_fac(ew, sig2, d2)) * vec[2];
    }
    if (il[1] != 0) {
      R d2 = dot(vec[1], vec[1]);
      bf += (-K / std::max((R)1 - d2 / R02, (R)0.01)) * vec[1];
      if (d2 < rm2) bf += lj_fac(ew, sig2, d2) * vec[1];
    }
  }
  bool bend = p->bend_all ? true : ((S.fl[i] & ORC_HET) != 0);
  if (bend) {
    R kb = (R)p->k_b;
    bf += kb * (il[1] * il[0] * vec[0] - cs[0] * vec[1] * il[1] * il[1]);
    bf += kb * (il[1] * il[2] * vec[2] - cs[1] * vec[1] * il[1] * il[1]);
    bf += kb * (-il[2] * il[1] * vec[1] + cs[1] * vec[2] * il[2] * il[2]);
    bf += kb * (-il[2] * il[3] * vec[3] + cs[2] * vec[2] * il[2] * il[2]);
  }
  return bf;
}

// Wall-particle vector (calc_cf<DST>, chrsim.cu:98-137): from the nearest
// wall point to the bead, sphere + opening + bleb. Returns false when not finite.
template <class R> bool wall_vec(const Geo<R> &g, V3<R> ri, bool sphere_only, V3<R> &vwp) {
  if (sphere_only) { // calc_cf<ICG>, chrsim.cu:155-163
    R d_r = len(ri);
    vwp = ri * -(g.R_n / d_r - (R)1);
  } else {
    V3<R> rb = {0, 0, g.nod + g.bod};
    V3<R> rr = ri - rb;
    if (ri.z < g.nod) {
      R d_r = len(ri);
      if (ri.z / d_r < g.noc) {
        vwp = ri * -(g.R_n / d_r - (R)1);
      } else {
        d_r = std::sqrt(ri.x * ri.x + ri.y * ri.y);
        vwp.x = -ri.x * (g.R_o / d_r - (R)1);
        vwp.y = -ri.y * (g.R_o / d_r - (R)1);
        vwp.z = ri.z - g.nod;
      }
    } else {
      R d_r = len(rr);
      if (-rr.z / d_r < g.boc) {
        vwp = rr * -(g.R_b / d_r - (R)1);
      } else {
        d_r = std::sqrt(rr.x * rr.x + rr.y * rr.y);
        vwp.x = -rr.x * (g.R_o / d_r - (R)1);
        vwp.y = -rr.y * (g.R_o / d_r - (R)1);
        vwp.z = rr.z + g.bod;
      }
    }
  }
  return std::isfinite(len(vwp));
}

// Confinement force on bead i.
template <class R> V3<R> wall_force(const Sys<R> &S, V3<R> ri, uint8_t fli) {
  const orc_params *p = S.p;
  V3<R> vwp;
  if (!wall_vec(S.g, ri, p->ff == ORC_FF_ICG, vwp)) return {0, 0, 0};
  if (p->ff != ORC_FF_LJ) { // calc_wpf, chrsim.cu:79-88 (Wang-Frenkel, eps = 1)
    R dwp = len(vwp);
    if (dwp > (R)p->rco) return {0, 0, 0};
    return ccp_wf((R)1, dwp) * vwp;
  }
  R d2 = dot(vwp, vwp);
  R sig2 = (R)(p->sigma * p->sigma), rm2 = (R)(std::pow(2.0, 1.0 / 3.0) * p->sigma * p->sigma);
  if (d2 < rm2) return lj_fac((R)p->eps_wall, sig2, d2) * vwp;
  if ((fli & ORC_LAD) && d2 < (R)(p->rc_lam * p->rc_lam)) return lj_fac((R)p->eps_lam, sig2, d2) * vwp;
  return {0, 0, 0};
}

// Pair force on i from j (non-bonded pair): calc_ppf<DST>/<ICG>
// (chrsim.cu:173-198) or the WCA/LJ set.
template <class R> inline V3<R> pair_force(const Sys<R> &S, V3<R> vpp, uint8_t fi, uint8_t fj) {
  const orc_params *p = S.p;
  bool hhi = (fi & ORC_HET) && (fj & ORC_HET);
  if (p->ff == ORC_FF_CCP) {
    R dpp = len(vpp);
    if (dpp > (R)p->aco) return {0, 0, 0};
    if (!hhi && dpp > (R)p->rco) return {0, 0, 0};
    return ccp_wf((R)p->e_p, dpp) * vpp;
  }
  if (p->ff == ORC_FF_ICG) {
    R dpp = len(vpp);
    if (dpp > (R)p->rco) return {0, 0, 0};
    return ((R)128 * ((R)3 * (R)p->rco - (R)3 * dpp)) * vpp;
  }
  R d2 = dot(vpp, vpp);
  R sig2 = (R)(p->sigma * p->sigma), rm2 = (R)(std::pow(2.0, 1.0 / 3.0) * p->sigma * p->sigma);
  if (d2 < rm2) return lj_fac((R)p->eps_wca, sig2, d2) * vpp;
  if (hhi && d2 < (R)(p->rc_lj * p->rc_lj)) return lj_fac((R)p->eps_het, sig2, d2) * vpp;
  return {0, 0, 0};
}

// LBS binding force (calc_lpf<DST>, chrsim.cu:201-212).
template <class R> inline V3<R> lbs_force(const Sys<R> &S, V3<R> vlp) {
  const orc_params *p = S.p;
  if (p->ff == ORC_FF_ICG) return {0, 0, 0};
  R dlp = len(vlp);
  R lco = (R)p->lco;
  if (dlp > lco) return {0, 0, 0};
  // ccp's arithmetic: nd and its powers in R, the rest in double
  R nd = dlp / lco;
  R d2 = nd * nd;
  const double lfac = (double)(R)p->e_l * (8.0 / (3.0 * (double)lco * (double)lco));
  return (R)(lfac * ((double)(d2 * d2 * d2) - 1.0)) * vlp;
}

// Sorted uniform grid over positions in R (sugrid.cu:15-77 semantics).
template <class R> struct Grid {
  std::vector<uint32_t> soi, beg, end;
};
// floor(v) as an int; values beyond +-2^30 and NaN (which only a blown-up
// state produces) saturate, so the conversion is always defined and such
// a bead lands outside every grid.
template <class R> inline int floor_index(R v) {
  const R lim = (R)(1 << 30);
  if (!(v > -lim)) return v != v ? (1 << 30) : -(1 << 30);
  if (v > lim) return 1 << 30;
  return (int)std::floor(v);
}
// Cell key in ccp's mixed int/uint arithmetic (calc_indexes, sugrid.cu:29-34):
// cps is unsigned there, so the sum wraps modulo 2^32.
template <class R> inline uint32_t cell_key(R csl, uint32_t cps, V3<R> r) {
  const int ix = floor_index(r.x / csl), iy = floor_index(r.y / csl), iz = floor_index(r.z / csl);
  const uint32_t iofst = (cps / 2) * (1 + cps + cps * cps);
  return iofst + (uint32_t)ix + (uint32_t)iy * cps + (uint32_t)iz * cps * cps;
}
template <class R> int build_grid(R csl, uint32_t cps, const V3<R> *r, int n, Grid<R> &G, uint32_t *uci_out = nullptr,
                                  uint32_t *sci_out = nullptr) {
  uint32_t n_c = cps * cps * cps;
  std::vector<uint32_t> uci(n), idx(n);
  for (int i = 0; i < n; ++i) {
    uci[i] = cell_key(csl, cps, r[i]);
    idx[i] = i;
  }
  std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return uci[a] < uci[b]; });
  G.soi = idx;
  G.beg.assign(n_c, 0);
  G.end.assign(n_c, 0);
  int bad = 0;
  for (int s = 0; s < n; ++s) { // find_cells_limits, sugrid.cu:53-77
    uint32_t cc = uci[idx[s]];
    if (cc >= n_c) { ++bad; continue; }
    if (s == 0) { G.beg[cc] = 0; }
    else {
      uint32_t cp = uci[idx[s - 1]];
      if (cp != cc) {
        G.beg[cc] = s;
        if (cp < n_c) G.end[cp] = s;
      }
    }
    if (s == n - 1) G.end[cc] = s + 1;
  }
  if (uci_out) std::copy(uci.begin(), uci.end(), uci_out);
  if (sci_out)
    for (int s = 0; s < n; ++s) sci_out[s] = uci[idx[s]];
  return bad;
}

// Short-range forces by ccp's 27-cell scan (calc_srf, chrsim.cu:277-317;
// calc_cell_srf :223-274): x outer, y, z inner; cell index in unsigned
// arithmetic, skipped only when >= n_c.
template <class R>
V3<R> short_range_cells(const Sys<R> &S, const V3<R> *r, const Grid<R> &PG, const Grid<R> &LG, int i) {
  V3<R> ri = r[i];
  const int ix = floor_index(ri.x / S.csl), iy = floor_index(ri.y / S.csl), iz = floor_index(ri.z / S.csl);
  uint32_t cps = S.cps;
  uint32_t iofst = (cps / 2) * (1 + cps + cps * cps);
  V3<R> srf = {0, 0, 0};
  uint8_t fi = S.fl[i];
  for (int nx = ix - 1; nx <= ix + 1; ++nx)
    for (int ny = iy - 1; ny <= iy + 1; ++ny)
      for (int nz = iz - 1; nz <= iz + 1; ++nz) {
        uint32_t nci = iofst + (uint32_t)nx + (uint32_t)ny * cps + (uint32_t)nz * cps * cps;
        if (nci >= S.n_c) continue;
        for (uint32_t s = PG.beg[nci]; s < PG.end[nci]; ++s) {
          int j = (int)PG.soi[s];
          if (std::abs(j - i) > 1) srf += pair_force(S, ri - r[j], fi, S.fl[j]);
        }
        if (fi & ORC_LAD)
          for (uint32_t s = LG.beg[nci]; s < LG.end[nci]; ++s) srf += lbs_force(S, ri - S.lbs[LG.soi[s]]);
      }
  return srf;
}

template <class R> V3<R> short_range_brute(const Sys<R> &S, const V3<R> *r, int i) {
  V3<R> ri = r[i];
  V3<R> srf = {0, 0, 0};
  uint8_t fi = S.fl[i];
  for (int j = 0; j < S.N; ++j)
    if (std::abs(j - i) > 1) srf += pair_force(S, ri - r[j], fi, S.fl[j]);
  if (fi & ORC_LAD)
    for (int l = 0; l < S.n_l; ++l) srf += lbs_force(S, ri - S.lbs[l]);
  return srf;
}

template <class R> void all_forces(const Sys<R> &S, const V3<R> *r, V3<R> *f, int method, int threads) {
  Grid<R> PG, LG;
  if (method == 0) {
    build_grid(S.csl, S.cps, r, S.N, PG);
    build_grid(S.csl, S.cps, S.lbs.data(), S.n_l, LG);
  }
#pragma omp parallel for schedule(static) num_threads(threads)
  for (int i = 0; i < S.N; ++i) {
    V3<R> fi = {0, 0, 0};
    fi += bonded(S, r, i);
    fi += wall_force(S, r[i], S.fl[i]);
    fi += (method == 0) ? short_range_cells(S, r, PG, LG, i) : short_range_brute(S, r, i);
    f[i] = fi;
  }
}

int nthreads(int t) {
#ifdef _OPENMP
  return t > 0 ? t : omp_get_max_threads();
#else
  (void)t;
  return 1;
#endif
}

// Optional integrator safeguard of the WCA/LJ set: rescale a bead's total
// force to magnitude fcap when it exceeds it (ccp itself has none).
template <class R> void cap_forces(std::vector<V3<R>> &f, R fcap) {
  for (auto &v : f) {
    R m = len(v);
    if (m > fcap) v = v * (fcap / m);
  }
}

template <class R>
int run_impl(const orc_params *p, const uint8_t *fl, const int32_t *co, const double *lbs, double *rio,
             uint64_t step0, int nsteps, int threads) {
  Sys<R> S = make_sys<R>(p, fl, co, lbs);
  int N = S.N;
  std::vector<V3<R>> r(N), er(N), f(N), ef(N), rn(N);
  for (int i = 0; i < N; ++i) r[i] = {(R)rio[3 * i], (R)rio[3 * i + 1], (R)rio[3 * i + 2]};
  R dt = (R)p->dt, sd = (R)p->sd;
  int T = nthreads(threads);
  for (int s = 0; s < nsteps; ++s) {
    uint64_t step = step0 + (uint64_t)s;
    // begin_iter (chrsim.cu:336-361)
#pragma omp parallel for schedule(static) num_threads(T)
    for (int i = 0; i < N; ++i) {
      if (p->sd == 0.0) { rn[i] = {0, 0, 0}; continue; }
      double z[3];
      noise_triple(p->seed, (uint32_t)i, step, z);
      rn[i] = {sd * (R)z[0], sd * (R)z[1], sd * (R)z[2]};
    }
    // exec_RK_1 (chrsim.cu:364-388)
    all_forces(S, r.data(), f.data(), 0, T);
    if (p->fcap > 0) cap_forces(f, (R)p->fcap);
    for (int i = 0; i < N; ++i) er[i] = r[i] + f[i] * dt + rn[i];
    // exec_RK_2 (chrsim.cu:391-416)
    all_forces(S, er.data(), ef.data(), 0, T);
    if (p->fcap > 0) cap_forces(ef, (R)p->fcap);
    for (int i = 0; i < N; ++i) r[i] = r[i] + ((R)0.5 * (ef[i] + f[i])) * dt + rn[i];
  }
  for (int i = 0; i < N; ++i) {
    rio[3 * i] = (double)r[i].x;
    rio[3 * i + 1] = (double)r[i].y;
    rio[3 * i + 2] = (double)r[i].z;
  }
  return 0;
}

template <class R>
int forces_impl(const orc_params *p, const uint8_t *fl, const int32_t *co, const double *lbs, const double *rin,
                double *fout, int method, int threads) {
  Sys<R> S = make_sys<R>(p, fl, co, lbs);
  std::vector<V3<R>> r(S.N), f(S.N);
  for (int i = 0; i < S.N; ++i) r[i] = {(R)rin[3 * i], (R)rin[3 * i + 1], (R)rin[3 * i + 2]};
  all_forces(S, r.data(), f.data(), method, nthreads(threads));
  for (int i = 0; i < S.N; ++i) {
    fout[3 * i] = (double)f[i].x;
    fout[3 * i + 1] = (double)f[i].y;
    fout[3 * i + 2] = (double)f[i].z;
  }
  return 0;
}

// ---------------------------------------------------------------- energies
// Potentials whose negative gradients are the forces above
// (Bibliography/forces.md): elastic 1/2 k_e (d - l_0)^2, bending
// k_b (1 - cos), Wang-Frenkel eps (1/d^2 - 1)(4/d^2 - 1)^2 (alpha = 1 for
// rc = 2), soft 128 (d + d_m/2)(d - d_m)^2, binding
// eps' (4/3 x^2 - 1/3 x^8 - 1), and the WCA/LJ set.
double wf_pot(double d2) { return (1.0 / d2 - 1.0) * (4.0 / d2 - 1.0) * (4.0 / d2 - 1.0); }
double lj_pot(double eps, double sig2, double d2) {
  double s6 = std::pow(sig2 / d2, 3.0);
  return 4.0 * eps * (s6 * s6 - s6);
}

} // namespace

extern "C" {

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  philox10(c, key[0], key[1]);
  std::memcpy(out, c, sizeof(c));
}

int orc_noise(uint64_t seed, uint32_t bead, uint64_t step, double z[3]) { return noise_triple(seed, bead, step, z); }

void orc_noise_batch(uint64_t seed, uint64_t step, int32_t n, double *z, int32_t *attempts) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) attempts[i] = noise_triple(seed, (uint32_t)i, step, z + 3 * (size_t)i);
}

int orc_cell_list(const float *r, int32_t n, float csl, int32_t cps, uint32_t *uci, uint32_t *sci, uint32_t *uoi,
                  uint32_t *soi, uint32_t *beg, uint32_t *end) {
  std::vector<V3<float>> rv(n);
  for (int i = 0; i < n; ++i) rv[i] = {r[3 * i], r[3 * i + 1], r[3 * i + 2]};
  Grid<float> G;
  int bad = build_grid<float>(csl, (uint32_t)cps, rv.data(), n, G, uci, sci);
  for (int i = 0; i < n; ++i) {
    uoi[i] = i;
    soi[i] = G.soi[i];
  }
  std::copy(G.beg.begin(), G.beg.end(), beg);
  std::copy(G.end.begin(), G.end.end(), end);
  return bad;
}

int orc_forces(const orc_params *p, const uint8_t *flags, const int32_t *chain_off, const double *lbs,
               const double *r, double *f, int32_t method, int32_t prec, int32_t threads) {
  if (prec == 32) return forces_impl<float>(p, flags, chain_off, lbs, r, f, method, threads);
  return forces_impl<double>(p, flags, chain_off, lbs, r, f, method, threads);
}

int orc_energy(const orc_params *p, const uint8_t *fl, const int32_t *co, const double *lbs, const double *r,
               double terms[5]) {
  Sys<double> S = make_sys<double>(p, fl, co, lbs);
  int N = S.N;
  auto R3 = [&](int i) { return V3<double>{r[3 * i], r[3 * i + 1], r[3 * i + 2]}; };
  std::vector<int32_t> chain(N);
  for (int c = 0; c < p->n_chains; ++c)
    for (int i = co[c]; i < co[c + 1]; ++i) chain[i] = c;
  double eb = 0, ea = 0, ep = 0, el = 0, ew = 0;
  double sig2 = p->sigma * p->sigma, rm2 = std::pow(2.0, 1.0 / 3.0) * sig2;
  // bonds
  for (int i = 0; i + 1 < N; ++i) {
    if (chain[i] != chain[i + 1]) continue;
    double d = len(R3(i + 1) - R3(i));
    if (p->bond_type == 0) eb += 0.5 * p->k_e * (d - p->l_0) * (d - p->l_0);
    else {
      double x = d * d / (p->fene_r0 * p->fene_r0);
      eb += -0.5 * p->fene_k * p->fene_r0 * p->fene_r0 * std::log(1.0 - x);
      if (d * d < rm2) eb += lj_pot(p->eps_wca, sig2, d * d) + p->eps_wca;
    }
  }
  // bending: angle at bead j between bonds (j-1,j) and (j,j+1)
  for (int j = 1; j + 1 < N; ++j) {
    if (chain[j - 1] != chain[j] || chain[j] != chain[j + 1]) continue;
    V3<double> b0 = R3(j) - R3(j - 1), b1 = R3(j + 1) - R3(j);
    double c = dot(b0, b1) / (len(b0) * len(b1));
    ea += p->k_b * (1.0 - c);
  }
  // pairs
  for (int i = 0; i < N; ++i)
    for (int j = i + 2; j < N; ++j) {
      double d2 = dot(R3(i) - R3(j), R3(i) - R3(j));
      bool hhi = (fl[i] & ORC_HET) && (fl[j] & ORC_HET);
      if (p->ff == ORC_FF_CCP) {
        double rc = hhi ? p->aco : p->rco;
        if (d2 < rc * rc) ep += p->e_p * (wf_pot(d2) - wf_pot(rc * rc));
      } else if (p->ff == ORC_FF_ICG) {
        double d = std::sqrt(d2), dm = p->rco;
        if (d < dm) ep += 128.0 * (d + dm / 2) * (d - dm) * (d - dm);
      } else {
        if (d2 < rm2) ep += lj_pot(p->eps_wca, sig2, d2) + p->eps_wca;
        if (hhi) {
          double rc2 = p->rc_lj * p->rc_lj;
          if (d2 < rm2) ep += -p->eps_het - lj_pot(p->eps_het, sig2, rc2);
          else if (d2 < rc2) ep += lj_pot(p->eps_het, sig2, d2) - lj_pot(p->eps_het, sig2, rc2);
        }
      }
    }
  // lbs
  if (p->ff != ORC_FF_ICG)
    for (int i = 0; i < N; ++i) {
      if (!(fl[i] & ORC_LAD)) continue;
      for (int l = 0; l < S.n_l; ++l) {
        double d = len(R3(i) - S.lbs[l]);
        if (d < p->lco) {
          double x2 = d * d / (p->lco * p->lco);
          el += p->e_l * (4.0 / 3.0 * x2 - 1.0 / 3.0 * x2 * x2 * x2 * x2 - 1.0);
        }
      }
    }
  // wall
  for (int i = 0; i < N; ++i) {
    V3<double> v;
    if (!wall_vec(S.g, R3(i), p->ff == ORC_FF_ICG, v)) continue;
    double d2 = dot(v, v);
    if (p->ff != ORC_FF_LJ) {
      if (d2 < p->rco * p->rco) ew += wf_pot(d2) - wf_pot(p->rco * p->rco);
    } else {
      if (d2 < rm2) ew += lj_pot(p->eps_wall, sig2, d2) + p->eps_wall;
      if (fl[i] & ORC_LAD) {
        double rc2 = p->rc_lam * p->rc_lam;
        if (d2 < rm2) ew += -p->eps_lam - lj_pot(p->eps_lam, sig2, rc2);
        else if (d2 < rc2) ew += lj_pot(p->eps_lam, sig2, d2) - lj_pot(p->eps_lam, sig2, rc2);
      }
    }
  }
  terms[0] = eb; terms[1] = ea; terms[2] = ep; terms[3] = el; terms[4] = ew;
  return 0;
}

int orc_run(const orc_params *p, const uint8_t *flags, const int32_t *chain_off, const double *lbs, double *r,
            uint64_t step0, int32_t nsteps, int32_t prec, int32_t threads) {
  if (prec == 32) return run_impl<float>(p, flags, chain_off, lbs, r, step0, nsteps, threads);
  return run_impl<double>(p, flags, chain_off, lbs, r, step0, nsteps, threads);
}

int orc_contacts(const float *r, int32_t n, float rc, int32_t bin_size, int32_t n_bins, uint32_t *map) {
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      float dx = r[3 * i] - r[3 * j], dy = r[3 * i + 1] - r[3 * j + 1], dz = r[3 * i + 2] - r[3 * j + 2];
      float d2 = (dx * dx + dy * dy) + dz * dz;
      if (sqrtf(d2) < rc) {
        uint32_t bi = (uint32_t)(i / bin_size), bj = (uint32_t)(j / bin_size);
        if (bi < (uint32_t)n_bins && bj < (uint32_t)n_bins) {
          map[(size_t)bi * n_bins + bj] += 1;
          map[(size_t)bj * n_bins + bi] += 1;
        }
      }
    }
  return 0;
}

int orc_contacts_grid(const float *r, int32_t n, float rc, int32_t bin_size, int32_t n_bins, uint32_t *map) {
  float lo[3] = {r[0], r[1], r[2]}, hi[3] = {r[0], r[1], r[2]};
  for (int i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], r[3 * i + a]);
      hi[a] = std::max(hi[a], r[3 * i + a]);
    }
  int dim[3];
  for (int a = 0; a < 3; ++a) dim[a] = std::max(1, (int)((hi[a] - lo[a]) / rc) + 1);
  auto cell = [&](int i, int a) { return std::min(dim[a] - 1, (int)((r[3 * i + a] - lo[a]) / rc)); };
  std::vector<std::vector<int>> cells((size_t)dim[0] * dim[1] * dim[2]);
  for (int i = 0; i < n; ++i) cells[((size_t)cell(i, 2) * dim[1] + cell(i, 1)) * dim[0] + cell(i, 0)].push_back(i);
  for (int i = 0; i < n; ++i) {
    int cx = cell(i, 0), cy = cell(i, 1), cz = cell(i, 2);
    for (int z = std::max(0, cz - 1); z <= std::min(dim[2] - 1, cz + 1); ++z)
      for (int y = std::max(0, cy - 1); y <= std::min(dim[1] - 1, cy + 1); ++y)
        for (int x = std::max(0, cx - 1); x <= std::min(dim[0] - 1, cx + 1); ++x)
          for (int j : cells[((size_t)z * dim[1] + y) * dim[0] + x]) {
            if (j <= i) continue;
            float dx = r[3 * i] - r[3 * j], dy = r[3 * i + 1] - r[3 * j + 1], dz = r[3 * i + 2] - r[3 * j + 2];
            float d2 = (dx * dx + dy * dy) + dz * dz;
            if (sqrtf(d2) < rc) {
              uint32_t bi = (uint32_t)(i / bin_size), bj = (uint32_t)(j / bin_size);
              if (bi < (uint32_t)n_bins && bj < (uint32_t)n_bins) {
                map[(size_t)bi * n_bins + bj] += 1;
                map[(size_t)bj * n_bins + bi] += 1;
              }
            }
          }
  }
  return 0;
}

int orc_observables(const double *r, const uint8_t *flags, int32_t n, double R_n, double nod, int32_t n_b,
                    double out[4], double *rcd) {
  auto R3 = [&](int i) { return V3<double>{r[3 * i], r[3 * i + 1], r[3 * i + 2]}; };
  V3<double> cm = {0, 0, 0};
  for (int i = 0; i < n; ++i) cm += R3(i);
  cm = cm * (1.0 / n);
  double rg2 = 0;
  for (int i = 0; i < n; ++i) rg2 += dot(R3(i) - cm, R3(i) - cm);
  double nop = 0;
  for (int i = 0; i + 1 < n; ++i) {
    V3<double> v = R3(i + 1) - R3(i);
    double c = dot(v, R3(i)) / (len(v) * len(R3(i)));
    nop += 0.5 * (3.0 * c * c - 1.0);
  }
  double ncf = 0;
  for (int i = 0; i < n; ++i) ncf += (r[3 * i + 2] < nod) ? 1.0 : 0.0;
  out[0] = len(cm);
  out[1] = rg2 / n;
  out[2] = n > 1 ? nop / (n - 1.0) : 0.0;
  out[3] = ncf / n;
  const double a = (double)1.154701f / 2.0, vb = 4.0 / 3.0 * M_PI * a * a * a;
  std::vector<double> rb(n_b);
  for (int k = 0; k < n_b; ++k) rb[k] = R_n * std::pow((k + 1.0) / n_b, 1.0 / 3.0);
  std::fill(rcd, rcd + 3 * n_b, 0.0);
  for (int i = 0; i < n; ++i) {
    double d = len(R3(i));
    int t = (flags[i] & ORC_HET) ? 0 : 1;
    for (int k = 0; k < n_b; ++k) {
      double v;
      if (d - a > rb[k]) v = 0;
      else if (d + a < rb[k]) v = vb;
      else {
        double ca = (d * d + a * a - rb[k] * rb[k]) / (2.0 * d * a);
        double cb = (rb[k] * rb[k] + d * d - a * a) / (2.0 * rb[k] * d);
        v = M_PI / 3.0 * a * a * a * (2.0 + ca) * (1.0 - ca) * (1.0 - ca) +
            M_PI / 3.0 * rb[k] * rb[k] * rb[k] * (2.0 + cb) * (1.0 - cb) * (1.0 - cb);
      }
      rcd[t * n_b + k] += v;
      rcd[2 * n_b + k] += v;
      if (k != n_b - 1) {
        rcd[t * n_b + k + 1] -= v;
        rcd[2 * n_b + k + 1] -= v;
      }
    }
  }
  const double shell = 4.0 / 3.0 * M_PI * R_n * R_n * R_n / n_b;
  for (int k = 0; k < 3 * n_b; ++k) rcd[k] /= shell;
  return 0;
}

int orc_sd_cp(const float *r, int32_t n, int32_t chromosome, float *sd, float *cp) {
  static const uint32_t lim[] = {0, 2284, 4599, 6687, 9338, 12508, 15193, 17477, 19792, 21880, 24531, 27701, 30386};
  const uint32_t N = (uint32_t)n, N_def = 30386;
  if (chromosome < 0 || chromosome >= 6 || (N != N_def && chromosome != 0)) return -1;
  const uint32_t lsdcp = (N == N_def) ? (lim[chromosome + 1] - lim[chromosome]) / 2 : N / 2;
  const uint32_t ncc = (N == N_def) ? 2 : 1;
  for (uint32_t ia = 0; ia < lsdcp; ++ia) {
    const uint32_t s = ia + 1;
    uint32_t i_s = lim[chromosome], i_e = (N == N_def) ? lim[chromosome + 1] : N;
    float a = 0.0f, c = 0.0f;
    for (uint32_t k = 0; k < ncc; ++k) {
      for (uint32_t p = i_s; p < i_e - s; ++p) {
        float dx = r[3 * (p + s)] - r[3 * p], dy = r[3 * (p + s) + 1] - r[3 * p + 1],
              dz = r[3 * (p + s) + 2] - r[3 * p + 2];
        float d = sqrtf((dx * dx + dy * dy) + dz * dz);
        a += d;
        if (d < 2.0f) c += 1000.0f;
      }
      i_s += N / ncc;
      i_e += N / ncc;
    }
    sd[ia] = a / (float)(ncc * (i_e - i_s - s));
    cp[ia] = c / (float)(ncc * (i_e - i_s - s));
  }
  return (int)lsdcp;
}

int orc_ccp_contact_map(const float *r, int32_t n, float *cm) {
  const uint32_t N = (uint32_t)n, N_def = 30386, px = 4;
  const float aco = 2.0f, cf = 1000.0f;
  uint32_t cms = (N == N_def) ? (N / 2) / px : N / px;
  uint32_t lcm = cms * (cms + 1) / 2;
  uint32_t ncc = (N == N_def) ? 2 : 1;
  for (uint32_t ia = 0; ia < lcm; ++ia) {
    float acc = 0.0f;
    // ccp's pixel coordinates, clamped to the map where the FP32 root of a
    // large index would round past the last row
    uint32_t ixp = std::min((uint32_t)(sqrtf((float)(2.0 * ia + 0.25)) - 0.5), cms - 1);
    uint32_t iyp = std::min((uint32_t)(ia - ixp * (ixp + 1.0) / 2.0), cms - 1);
    uint32_t sx = ixp * px, ex = sx + px, sy = iyp * px, ey = sy + px;
    for (uint32_t c = 0; c < ncc; ++c) {
      for (uint32_t a = sx; a < ex; ++a)
        for (uint32_t b = sy; b < ey; ++b) {
          float dx = r[3 * a] - r[3 * b], dy = r[3 * a + 1] - r[3 * b + 1], dz = r[3 * a + 2] - r[3 * b + 2];
          float d = sqrtf(dx * dx + dy * dy + dz * dz);
          if (d < aco) acc += cf;
        }
      sx += N / ncc; ex += N / ncc; sy += N / ncc; ey += N / ncc;
    }
    cm[ia] = acc / (float)(ncc * px * px);
  }
  return 0;
}

} // extern "C"
