/*
 * CPU oracle for the ccp Brownian-dynamics step — TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product (libccpmi) never
 * links, calls or falls back to it.
 *
 * It is a scalar FP64 (and FP32) restatement of the algorithm in the
 * reference program `ccp` (Marco-Mendivil-Carboni/TFM, Program/src). Its
 * FP32 mode keeps ccp's mixed widths: the Wang-Frenkel and binding factors
 * are evaluated in FP64 where ccp's double literals promote them
 * (chrsim.cu:87,184,211):
 *   - bonded, bending, wall, pair and lamina-binding forces: chrsim.cu:13-317
 *   - stochastic Heun step:                                   chrsim.cu:336-416
 *   - sorted uniform grid (cell list):                        sugrid.cu:15-133
 *   - contact map (packed lower triangle):                    chrana.cu:48-84
 * plus the build's own Philox4x32-10 counter RNG (Salmon, Moraes, Dror,
 * Shaw, SC'11) and Box–Muller normals with ccp's 5-sigma redraw rule
 * (chrsim.cu:348-356), and the optional WCA/LJ/FENE force field.
 *
 * How each part is pinned (see tests/test_oracle_*.py):
 *   Philox          Random123 known-answer vectors + rocRAND's host engine
 *   normals         moments, redraw closed form
 *   cell list       brute-force std::stable_sort comparison, ccp limits rule
 *   forces          finite differences of the potentials (forces.md closed
 *                   forms), cutoff zeros, brute-force O(N^2) pair sums
 *   integrator      free diffusion variance 2 kT dt, harmonic dimer
 *   contacts        brute-force O(N^2) counts
 * The ccp bending term for mixed A/B chains is not a gradient: it is pinned
 * as the gradient of the all-angle bending energy masked to the A beads,
 * 1[i is A] (-grad V_all), by finite differences.
 */
#ifndef CCP_ORACLE_H
#define CCP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Force-field selector. */
enum orc_ff {
  ORC_FF_CCP = 0, /* ccp default simulation (DST): Wang-Frenkel pair + wall, LBS binding */
  ORC_FF_ICG = 1, /* ccp initial-condition generation: soft repulsion, sphere wall */
  ORC_FF_LJ = 2   /* optional set: WCA everywhere, LJ tail het-het, wall WCA + lamina tail */
};

/* Per-bead flag bits. A ccp "LADh" (A) bead is HET|LAD, "LNDe" (B) is 0. */
enum { ORC_HET = 1, ORC_LAD = 2 };

/* Resolved parameters (plain C layout, mirrored by oracle/pyoracle.py). */
typedef struct orc_params {
  int32_t N;        /* beads */
  int32_t n_l;      /* lamina binding sites */
  int32_t n_chains; /* chains; chain_off has n_chains+1 entries */
  int32_t ff;       /* enum orc_ff */
  int32_t bend_all; /* 0: bending only on HET beads (ccp), 1: all beads */
  int32_t bond_type;/* 0 harmonic (ccp), 1 FENE + bonded WCA */
  int32_t cps;      /* cells per side of the cell grid */
  int32_t pad0;
  double R_n, R_o, R_b; /* nucleus, opening and bleb radii */
  double csl;           /* cell side length */
  double k_e, l_0, k_b; /* bond and bending constants */
  double e_p, rco, aco; /* Wang-Frenkel pair energy, repulsive and attractive cutoffs */
  double e_l, lco;      /* LBS binding energy and cutoff */
  double fene_k, fene_r0;
  double eps_wca, sigma, eps_het, rc_lj; /* WCA/LJ pair set */
  double eps_wall, eps_lam, rc_lam;      /* wall repulsion, lamina attraction */
  double dt, sd;        /* timestep, noise standard deviation */
  uint64_t seed;        /* Philox key */
  double fcap;          /* integrator force cap |F| <= fcap per bead and stage (0 = off) */
} orc_params;

/* Philox4x32-10 bijection: out = philox(ctr, key). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Gaussian noise triple for one bead and step, in units of one standard
 * deviation; returns the number of redraws. */
int orc_noise(uint64_t seed, uint32_t bead, uint64_t step, double z[3]);

/* orc_noise for beads [0, n) of one step: z is n x 3, attempts n. */
void orc_noise_batch(uint64_t seed, uint64_t step, int32_t n, double *z, int32_t *attempts);

/* ccp's sorted uniform grid for N positions (float, xyz interleaved):
 * uci/uoi unsorted keys/indices, sci/soi stable-sorted, beg/end per cell
 * (empty cells 0/0). Returns the number of objects whose key is out of range. */
int orc_cell_list(const float *r, int32_t n, float csl, int32_t cps,
                  uint32_t *uci, uint32_t *sci, uint32_t *uoi, uint32_t *soi,
                  uint32_t *beg, uint32_t *end);

/* Total force on every bead (xyz interleaved). method 0: ccp cell scan,
 * 1: brute force over all pairs. prec 64 or 32. threads <= 0 means all. */
int orc_forces(const orc_params *p, const uint8_t *flags, const int32_t *chain_off,
               const double *lbs, const double *r, double *f, int32_t method,
               int32_t prec, int32_t threads);

/* Potential energy split into terms: [0] bonds, [1] bending, [2] pairs,
 * [3] LBS binding, [4] wall. Defined only where the force is a gradient. */
int orc_energy(const orc_params *p, const uint8_t *flags, const int32_t *chain_off,
               const double *lbs, const double *r, double terms[5]);

/* nsteps stochastic-Heun steps starting at step counter step0, in place. */
int orc_run(const orc_params *p, const uint8_t *flags, const int32_t *chain_off,
            const double *lbs, double *r, uint64_t step0, int32_t nsteps,
            int32_t prec, int32_t threads);

/* Binned contact counts of ordered pairs i != j with |r_i - r_j| < rc,
 * accumulated into map[bin(i) * n_bins + bin(j)], bin(i) = i / bin_size.
 * Distances use FP32 with no fused operations, and ccp's contact test
 * (calc_cm, chrana.cu:73-76): sqrtf((dx*dx + dy*dy) + dz*dz) < rc. */
int orc_contacts(const float *r, int32_t n, float rc, int32_t bin_size,
                 int32_t n_bins, uint32_t *map);

/* Same counts as orc_contacts, found through a hash grid of cell size
 * >= rc (for large n); the distance test is identical. */
int orc_contacts_grid(const float *r, int32_t n, float rc, int32_t bin_size,
                      int32_t n_bins, uint32_t *map);

/* ccp's per-frame contact map (chrana.cu:48-84): packed lower triangle of
 * cms*(cms+1)/2 floats, 4-bead pixels, factor 1000, cutoff aco = 2, with the
 * two-copy average when n == 30386. */
int orc_ccp_contact_map(const float *r, int32_t n, float *cm);

/* ccp's per-frame observables (chrana.cu:372-480) in FP64 for n beads:
 * out[0] centre-of-mass distance, [1] radius of gyration squared,
 * [2] nematic order parameter, [3] fraction with z < nod; rcd (3 x n_b) the
 * radial chromatin density of A beads (HET flag), B beads and all beads
 * with the lens-volume formula of ccp. */
int orc_observables(const double *r, const uint8_t *flags, int32_t n, double R_n, double nod, int32_t n_b,
                    double out[4], double *rcd);

/* ccp's mean distance d(s) and contact probability P(s) x 1000 of one
 * chromosome, s = 1..lsdcp (calc_sd_cp, chrana.cu:13-45, ranges of
 * chrana.cu:93-97,484-491): FP32 sums over p of |r[p+s] - r[p]| and of 1000
 * per d < 2, both copies when n = 30386 (chromosome 0..5 of CHR_LIM, lsdcp =
 * half its length), else chromosome 0 = beads [0, n), lsdcp = n/2.
 * Returns lsdcp, or -1 for a chromosome that does not exist. */
int orc_sd_cp(const float *r, int32_t n, int32_t chromosome, float *sd, float *cp);

#ifdef __cplusplus
}
#endif

#endif
