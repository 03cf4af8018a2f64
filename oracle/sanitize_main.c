/* Test infrastructure only: a driver that runs the oracle under
 * AddressSanitizer and UndefinedBehaviorSanitizer (make -C oracle sanitize).
 * Every entry point the tests use, on a few molecules, both Fock modes; it
 * prints one line per molecule and exits non-zero when the two Fock modes'
 * SCF energies differ by more than 1e-8 eV or an entry point fails (the
 * sanitizers abort on their own findings). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "mndo_oracle.h"

typedef struct {
  const char* name;
  int n;
  int z[12];
  double xyz[36];
} mol_t;

static int run(const mol_t* m) {
  const int n = m->n;
  const int nao = or_n_ao(n, m->z);
  const long nb = or_pair_block_count(n, m->z);
  const int T = nao * (nao + 1) / 2;
  double* blocks = malloc(sizeof(double) * (size_t)nb);
  double* h = malloc(sizeof(double) * T);
  double* p = malloc(sizeof(double) * T);
  double* f1 = malloc(sizeof(double) * T);
  double* f2 = malloc(sizeof(double) * T);
  double* a = malloc(sizeof(double) * nao * nao);
  double* w = malloc(sizeof(double) * nao);
  double* en = malloc(sizeof(double) * 101);
  int bad = 0;
  bad |= or_pair_blocks(n, m->z, m->xyz, blocks) != 0;
  bad |= or_hcore(n, m->z, m->xyz, blocks, h) != 0;
  bad |= or_guess_density(n, m->z, p) != 0;
  bad |= or_fock_ref(n, m->z, blocks, h, p, f1) != 0;
  bad |= or_fock_nddo(n, m->z, blocks, h, p, f2) != 0;
  double df = 0.0;
  for (int i = 0; i < T; ++i) df = fmax(df, fabs(f1[i] - f2[i]));
  for (int i = 0; i < nao; ++i)
    for (int j = 0; j <= i; ++j) a[i * nao + j] = a[j * nao + i] = f1[i * (i + 1) / 2 + j];
  bad |= or_eigh(nao, a, w) != 0;
  const double e0 = or_energy(nao, h, f1, p);
  const double ecc = or_core_core(n, m->z, m->xyz, blocks);
  (void)or_get_value(n, m->z, blocks, 0, 0, nao - 1, nao - 1);
  (void)or_overlap(n, m->z, m->xyz, 0, nao - 1);
  or_scf_result r1, r2;
  bad |= or_scf(n, m->z, m->xyz, 1e-6, 100, 0, &r1, en, NULL, NULL, NULL) != 0;
  bad |= or_scf(n, m->z, m->xyz, 1e-6, 100, 1, &r2, NULL, p, f2, w) != 0;
  const double de = fabs(r1.e_tot - r2.e_tot);
  printf("%-12s N %3d  E0 %.6f  Ecc %.6f  Etot %.10f / %.10f  |dF| %.1e  steps %d/%d\n", m->name,
         nao, e0, ecc, r1.e_tot, r2.e_tot, df, r1.steps, r2.steps);
  bad |= !(r1.converged && r2.converged) || de > 1e-8 || df > 1e-10 * fmax(1.0, fabs(f1[0]));
  free(blocks); free(h); free(p); free(f1); free(f2); free(a); free(w); free(en);
  return bad;
}

int main(void) {
  const mol_t mols[] = {
      {"H2", 2, {1, 1}, {0, 0, 0, 0, 0, 0.74}},
      {"water", 3, {8, 1, 1}, {0, 0, 0, 0.7571, 0.5861, 0, -0.7571, 0.5861, 0}},
      {"formaldehyde", 4, {6, 8, 1, 1}, {0, 0, 0, 0, 0, 1.216, 0.94, 0, -0.55, -0.94, 0, -0.55}},
      {"methanimine", 5, {6, 7, 1, 1, 1},
       {0, 0, 0, 1.27, 0, 0, -0.55, 0.94, 0, -0.55, -0.94, 0, 1.8, 0.87, 0}},
      {"ethane", 8, {6, 6, 1, 1, 1, 1, 1, 1},
       {0, 0, 0.765, 0, 0, -0.765, 1.02, 0, 1.16, -0.51, 0.88, 1.16, -0.51, -0.88, 1.16,
        -1.02, 0, -1.16, 0.51, -0.88, -1.16, 0.51, 0.88, -1.16}},
  };
  int bad = 0;
  for (size_t i = 0; i < sizeof(mols) / sizeof(mols[0]); ++i) bad |= run(&mols[i]);
  printf(bad ? "sanitize: FAIL\n" : "sanitize: ok\n");
  return bad;
}
