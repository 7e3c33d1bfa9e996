/* Host-only driver for an AddressSanitizer / UBSan build of the oracle
 * (make -C oracle _build/oracle_sanity): every oracle entry point once on a
 * small confined chain with LBS sites, so out-of-bounds indexing in the
 * restatement shows up on the CPU. Test infrastructure only. */
#include "ccp_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NB 300
#define NL 12

static double urand(unsigned *s) {
  *s = *s * 1664525u + 1013904223u;
  return ((*s >> 8) + 0.5) / 16777216.0;
}

int main(void) {
  static double r[3 * NB], f[3 * NB], lbs[3 * NL];
  static float rf[3 * NB];
  static uint8_t flags[NB];
  int32_t chain_off[3] = {0, 140, NB};
  unsigned s = 12345u;
  const double R_n = 8.0;
  /* confined random walk with unit steps */
  for (int i = 0; i < NB; ++i) {
    for (;;) {
      double v[3], n = 0.0;
      for (int a = 0; a < 3; ++a) {
        v[a] = urand(&s) - 0.5;
        n += v[a] * v[a];
      }
      n = sqrt(n);
      double p[3];
      for (int a = 0; a < 3; ++a) p[a] = (i == 0 || i == 140 ? 0.0 : r[3 * (i - 1) + a]) + v[a] / n;
      if (sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) < R_n - 1.0) {
        memcpy(&r[3 * i], p, sizeof(p));
        break;
      }
    }
    flags[i] = (uint8_t)((i / 10) % 2 ? ORC_HET : 0);
  }
  for (int l = 0; l < NL; ++l) {
    double v[3] = {urand(&s) - 0.5, urand(&s) - 0.5, urand(&s) - 0.5};
    double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int a = 0; a < 3; ++a) lbs[3 * l + a] = (R_n - 1.154701) * v[a] / n;
  }
  orc_params p;
  memset(&p, 0, sizeof(p));
  p.N = NB;
  p.n_l = NL;
  p.n_chains = 2;
  p.ff = ORC_FF_CCP;
  p.R_n = R_n;
  p.csl = 2.0;
  p.cps = 2 * (int)ceil(R_n / 2.0);
  p.k_e = 128.0; p.l_0 = 1.0; p.k_b = 1.0;
  p.e_p = 0.5; p.rco = 1.154701; p.aco = 2.0;
  p.e_l = 8.0; p.lco = 0.5;
  p.dt = 1.0 / 2048; p.sd = sqrt(2.0 * 0.003356 * 298.0 / 2048);
  p.seed = 7;
  int rc = 0;
  /* soft-repulsion relaxation without noise first, as the initial condition
   * does: the Wang-Frenkel force diverges for overlapping beads */
  p.ff = ORC_FF_ICG;
  const double sd0 = p.sd;
  p.sd = 0.0;
  rc |= orc_run(&p, flags, chain_off, lbs, r, 0, 2000, 64, 1);
  p.ff = ORC_FF_CCP;
  p.sd = sd0;
  rc |= orc_forces(&p, flags, chain_off, lbs, r, f, 0, 64, 1);
  rc |= orc_forces(&p, flags, chain_off, lbs, r, f, 1, 32, 1);
  double e[5];
  rc |= orc_energy(&p, flags, chain_off, lbs, r, e);
  rc |= orc_run(&p, flags, chain_off, lbs, r, 0, 20, 64, 1);
  for (int i = 0; i < 3 * NB; ++i) rf[i] = (float)r[i];
  static uint32_t uci[NB], sci[NB], uoi[NB], soi[NB], beg[8 * 8 * 8], end[8 * 8 * 8];
  rc |= orc_cell_list(rf, NB, 2.0f, p.cps, uci, sci, uoi, soi, beg, end);
  static uint32_t map[30 * 30];
  rc |= orc_contacts(rf, NB, 1.5f, 10, 30, map);
  rc |= orc_contacts_grid(rf, NB, 1.5f, 10, 30, map);
  static float cm[75 * 76 / 2];
  rc |= orc_ccp_contact_map(rf, NB, cm);
  double obs[4], rcd[3 * 16];
  rc |= orc_observables(r, flags, NB, R_n, R_n, 16, obs, rcd);
  static float sd[NB / 2], cp[NB / 2];
  rc |= orc_sd_cp(rf, NB, 0, sd, cp) < 0;
  double z[3];
  orc_noise(7, 3, 11, z);
  /* a blown-up state (NaN and huge coordinates) must stay defined behaviour:
   * such beads fall outside every grid; the results themselves are ignored */
  static double rb[3 * NB];
  memcpy(rb, r, sizeof(rb));
  rb[3 * 5] = NAN;
  rb[3 * 6 + 1] = 1e30;
  rb[3 * 7 + 2] = -1e30;
  for (int i = 0; i < 3 * NB; ++i) rf[i] = (float)rb[i];
  (void)orc_cell_list(rf, NB, 2.0f, p.cps, uci, sci, uoi, soi, beg, end);
  (void)orc_forces(&p, flags, chain_off, lbs, rb, f, 0, 64, 1);
  int finite = 1;
  for (int i = 0; i < 3 * NB; ++i) finite &= isfinite(r[i]) != 0;
  printf("oracle sanity: rc=%d finite=%d rg2=%.4f\n", rc, finite, obs[1]);
  return rc != 0 || !finite;
}
