/*
 * nr_check.c -- host replay of the level-synchronous indexx of the radial
 * sort (ahf_amd/csrc/unbind_gpu.hip, the tie-order kernels) against the
 * oracle's sequential indexx (oracle/ahf_oracle.c, itself pinned against the
 * reference's general.c:1093-1153).  Several arrays are laid end to end as
 * the kernels lay out the haloes with ties; every level partitions every
 * segment at once through the same header rules (ahf_amd/csrc/nr_order.h),
 * with the P and Q lists built from prefix counts as the kernels build them.
 * Inputs are tie-heavy: a few distinct values, all equal, sorted, reversed,
 * organ-pipe, lattice-shell distances, random with duplicates.  Any
 * permutation that differs fails.
 *
 * usage: nr_check [arrays] [seed]      (links oracle/liboracle.so)
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../ahf_amd/csrc/nr_order.h"

void oracle_indexx(uint64_t n, const double *arr1, uint64_t *indx1);

static uint64_t rng_state;
static uint64_t rnd(void)
{
  uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static void fill(double *a, int64_t n, int kind)
{
  for (int64_t i = 0; i < n; i++) {
    switch (kind) {
    case 0: a[i] = (double)(rnd() % 3); break;                    /* three values */
    case 1: a[i] = 1.0; break;                                    /* all equal */
    case 2: a[i] = (double)(i / 3); break;                        /* sorted, runs of 3 */
    case 3: a[i] = (double)((n - i) / 2); break;                  /* reversed, runs of 2 */
    case 4: a[i] = (double)(i < n / 2 ? i : n - i); break;        /* organ pipe */
    case 5: {                                                     /* lattice shells */
      const int64_t x = (int64_t)(rnd() % 9) - 4, y = (int64_t)(rnd() % 9) - 4, z = (int64_t)(rnd() % 9) - 4;
      a[i] = (double)(x * x + y * y + z * z) * 0.015625;
      break;
    }
    case 6: a[i] = (double)(rnd() % (uint64_t)(n / 4 + 1)); break; /* many duplicates */
    default: a[i] = (double)(rnd() >> 11) * 0x1p-53; break;        /* distinct */
    }
  }
}

/* the kernels' driver, on the host: one array per "halo" at base[h] */
static int replay(const double *key, uint32_t *idx, int64_t m, int nh, const int64_t *base, int *levels)
{
  int64_t cap = m / 2 + nh + 1;
  int64_t *l = malloc(cap * 8), *ir = malloc(cap * 8), *b = malloc(cap * 8);
  int64_t *l2 = malloc(cap * 8), *ir2 = malloc(cap * 8), *b2 = malloc(cap * 8);
  int64_t *K = malloc(cap * 8), *I = malloc(cap * 8), *J = malloc(cap * 8), *coff = malloc((cap + 1) * 8);
  double *a = malloc(cap * 8);
  uint32_t *pv = malloc(cap * 4), *cur = malloc(cap * 4);
  uint8_t *act = malloc(cap);
  int64_t *seg = malloc(m * 8), *sGE = malloc((m + 1) * 8), *sLE = malloc((m + 1) * 8);
  uint8_t *fl = malloc(m);
  uint32_t *P = malloc(m * 4), *Q = malloc(m * 4);
  int64_t S = 0;
  for (int h = 0; h < nh; h++) {
    if (base[h + 1] - base[h] < 2) continue;
    l[S] = base[h];
    ir[S] = base[h + 1] - 1;
    b[S] = base[h];
    for (int64_t g = base[h]; g < base[h + 1]; g++) seg[g] = S;
    S++;
  }
  for (int h = 0; h < nh; h++)
    if (base[h + 1] - base[h] < 2)
      for (int64_t g = base[h]; g < base[h + 1]; g++) seg[g] = -1;
  int lev = 0, rc = 0;
  while (S > 0) {
    lev++;
    /* per segment: insertion sort or median of three */
    for (int64_t s = 0; s < S; s++) {
      if (ir[s] - l[s] < NR_SMALL) {
        nr_insertion(key, idx, b[s], l[s], ir[s]);
        act[s] = 0;
      } else {
        act[s] = 1;
        a[s] = nr_median3(key, idx, b[s], l[s], ir[s], &pv[s]);
      }
    }
    /* per element: P and Q membership, prefix counts, list slots */
    for (int64_t g = 0; g < m; g++) {
      const int64_t s = seg[g];
      fl[g] = (s >= 0 && act[s]) ? (uint8_t)nr_flags(nr_key(key, idx, b[s], g), a[s], g, l[s], ir[s]) : 0;
    }
    sGE[0] = sLE[0] = 0;
    for (int64_t g = 0; g < m; g++) {
      sGE[g + 1] = sGE[g] + (fl[g] & 1);
      sLE[g + 1] = sLE[g] + ((fl[g] >> 1) & 1);
    }
    for (int64_t g = 0; g < m; g++) {
      const int64_t s = seg[g];
      if (s < 0 || !act[s]) continue;
      if (fl[g] & 1) P[l[s] + (sGE[g] - sGE[l[s] + 2])] = (uint32_t)g;
      if (fl[g] & 2) Q[l[s] + (sLE[ir[s]] - sLE[g + 1])] = (uint32_t)g;
    }
    /* per segment: where the scans meet */
    for (int64_t s = 0; s < S; s++) {
      if (!act[s]) continue;
      const int64_t nP = sGE[ir[s] + 1] - sGE[l[s] + 2], nQ = sLE[ir[s]] - sLE[l[s]];
      int js;
      if (nr_meet(P + l[s], Q + l[s], nP, nQ, &K[s], &I[s], &J[s], &js)) { rc = -1; goto out; }
      cur[s] = js ? idx[Q[l[s] + K[s] - 1]] : idx[J[s]];
    }
    /* per element: the swaps */
    for (int64_t g = 0; g < m; g++) {
      const int64_t s = seg[g];
      if (s < 0 || !act[s] || g - l[s] >= K[s]) continue;
      const uint32_t p = P[g], q = Q[g], t = idx[p];
      idx[p] = idx[q];
      idx[q] = t;
    }
    /* per segment: the pivot, and the next segments */
    coff[0] = 0;
    for (int64_t s = 0; s < S; s++) {
      int64_t c = 0;
      if (act[s]) {
        idx[l[s]] = cur[s];
        idx[J[s]] = pv[s];
        c = (J[s] - 1 - l[s] >= 1) + (ir[s] - I[s] >= 1);
      }
      coff[s + 1] = coff[s] + c;
    }
    for (int64_t s = 0; s < S; s++) {
      if (!act[s]) continue;
      int64_t c = coff[s];
      if (J[s] - 1 - l[s] >= 1) { l2[c] = l[s]; ir2[c] = J[s] - 1; b2[c] = b[s]; c++; }
      if (ir[s] - I[s] >= 1) { l2[c] = I[s]; ir2[c] = ir[s]; b2[c] = b[s]; }
    }
    for (int64_t g = 0; g < m; g++) {
      const int64_t s = seg[g];
      if (s < 0) continue;
      if (!act[s]) { seg[g] = -1; continue; }
      const int left = J[s] - 1 - l[s] >= 1, right = ir[s] - I[s] >= 1;
      if (g < J[s]) seg[g] = left ? coff[s] : -1;
      else if (g >= I[s]) seg[g] = right ? coff[s] + left : -1;
      else seg[g] = -1;
    }
    S = coff[S];
    int64_t *t;
    t = l; l = l2; l2 = t;
    t = ir; ir = ir2; ir2 = t;
    t = b; b = b2; b2 = t;
  }
out:
  *levels = lev;
  free(l); free(ir); free(b); free(l2); free(ir2); free(b2); free(K); free(I); free(J); free(coff);
  free(a); free(pv); free(cur); free(act); free(seg); free(sGE); free(sLE); free(fl); free(P); free(Q);
  return rc;
}

int main(int argc, char **argv)
{
  const int runs = argc > 1 ? atoi(argv[1]) : 2000;
  rng_state = argc > 2 ? strtoull(argv[2], NULL, 10) : 1;
  long bad = 0, arrays = 0, maxlev = 0;
  for (int r = 0; r < runs; r++) {
    const int nh = 1 + (int)(rnd() % 6);
    int64_t base[8];
    base[0] = 0;
    for (int h = 0; h < nh; h++) {
      int64_t n = (int64_t)(rnd() % (r % 10 == 0 ? 5000 : 300));
      if (rnd() % 8 == 0) n = (int64_t)(rnd() % 12);
      base[h + 1] = base[h] + n;
    }
    const int64_t m = base[nh];
    double *key = malloc((m + 1) * 8);
    uint32_t *idx = malloc((m + 1) * 4);
    for (int h = 0; h < nh; h++) {
      fill(key + base[h], base[h + 1] - base[h], (int)(rnd() % 8));
      for (int64_t g = base[h]; g < base[h + 1]; g++) idx[g] = (uint32_t)(g - base[h] + 1);
    }
    int lev = 0;
    if (replay(key, idx, m, nh, base, &lev)) {
      printf("run %d: a list ran short\n", r);
      bad++;
    }
    if (lev > maxlev) maxlev = lev;
    for (int h = 0; h < nh; h++) {
      const int64_t n = base[h + 1] - base[h];
      double *a1 = malloc((n + 1) * 8);
      uint64_t *ix = malloc((n + 1) * 8);
      for (int64_t i = 0; i < n; i++) a1[i + 1] = key[base[h] + i];
      oracle_indexx((uint64_t)n, a1, ix);
      for (int64_t i = 0; i < n; i++)
        if ((uint64_t)idx[base[h] + i] != ix[i + 1]) {
          if (bad < 5) printf("run %d halo %d (n %lld): position %lld differs\n", r, h, (long long)n, (long long)i);
          bad++;
          break;
        }
      arrays++;
      free(a1);
      free(ix);
    }
    free(key);
    free(idx);
  }
  printf("{\"arrays\": %ld, \"mismatches\": %ld, \"max_levels\": %ld}\n", arrays, bad, maxlev);
  return bad != 0;
}
