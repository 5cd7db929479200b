/*
 * fold64_check.c -- test-only host replay of an FP64 running-sum fold
 * (fold64.h; a wavefront's loop, one array slot per lane), against the
 * sequential sum d = d + t with FP64 rounding, on random and adversarial
 * runs: exact ties on the running value's grid, binade crossings in both
 * directions, sign changes, zeros, negative zeros, subnormals, huge and
 * non-finite-free extreme terms.  Any running value that differs in a bit
 * counts as a mismatch.
 *
 *   fold64_check <runs> <seed>     prints {"runs": R, "mismatches": M, ...}
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fold64.h"

static uint64_t rng_s;
static uint64_t rnd(void)
{
  rng_s ^= rng_s << 13;
  rng_s ^= rng_s >> 7;
  rng_s ^= rng_s << 17;
  return rng_s;
}
static double unif(void) { return (double)(rnd() >> 11) * (1.0 / 9007199254740992.0); }

static int64_t rounds_total, plain_total;

/* the wavefront loop: m <= 64 terms onto *d, out[l] the running value after
 * term l; lanes are array slots, ballots and scans are loops */
static void fold64_emu(double *d, const double *t, int m, double *out)
{
  int start = 0;
  while (start < m) {
    rounds_total++;
    int e, neg;
    int64_t K0;
    if (!ahf_f64_binade(*d, &e, &K0, &neg)) {
      *d = *d + t[start];
      out[start] = *d;
      start++;
      plain_total++;
      continue;
    }
    int64_t q[64], r[64];
    int lok[64], up[64], tie[64];
    for (int l = start; l < m; l++) {
      lok[l] = ahf_f64_lane(t[l], e, neg, &q[l], &up[l], &tie[l]);
      r[l] = q[l] + up[l];
    }
    /* ties in lane order: the even one of K + q and K + q + 1, K the value
     * before the lane (the sum of the lanes below it) */
    for (int l = start; l < m; l++) {
      if (!(lok[l] && tie[l])) continue;
      int64_t below = 0;
      for (int k = start; k < l; k++) below += lok[k] ? r[k] : 0;
      r[l] = ((K0 + below + q[l]) & 1) ? q[l] + 1 : q[l];
    }
    int64_t P = 0;
    int bad = m;
    for (int l = start; l < m; l++) {
      const int64_t rl = lok[l] ? r[l] : 0;
      const int64_t Kb = K0 + P;
      P += rl;
      if (!(lok[l] && ahf_f64_in(Kb, q[l]))) {
        bad = l;
        break;
      }
      out[l] = ahf_f64_make(K0 + P, e, neg);
    }
    if (bad > start) *d = out[bad - 1];
    if (bad < m) {
      *d = *d + t[bad];
      out[bad] = *d;
      start = bad + 1;
      plain_total++;
    } else {
      start = m;
    }
  }
}

/* one term of a kind, given the running value d */
static double term(double d, int kind)
{
  const double ad = fabs(d);
  int e = 0;
  if (ad > 0) (void)frexp(ad, &e);
  const double ulp = ad > 0 ? ldexp(1.0, e - 53) : 0x1p-1074;
  const double s = (rnd() & 1) ? 1.0 : -1.0;
  switch (kind) {
  case 0: /* random magnitude */
    return s * ldexp(0.5 + unif(), (int)(rnd() % 80) - 60);
  case 1: /* exact half-ulp ties and near ties */
    return s * ((double)(int64_t)(rnd() % 64) + 0.5) * ulp * (rnd() % 4 == 0 ? 1.0 : 2.0);
  case 2: /* terms of the running value's size: binade crossings */
    return s * ad * (0.25 + 2.0 * unif());
  case 3: /* small terms of many bits */
    return s * ulp * (1.0 + unif() * 1000.0);
  case 4: /* zeros, negative zeros */
    return (rnd() & 1) ? 0.0 : -0.0;
  case 5: /* subnormals */
    return s * ldexp((double)(rnd() % 1000 + 1), -1074);
  case 6: /* huge */
    return s * ldexp(0.5 + unif(), 1000 + (int)(rnd() % 20));
  case 7: /* the profiles' kind: positive terms of one scale */
    return ldexp(unif(), -20);
  case 8: /* a random walk of one scale */
    return s * unif();
  default: { /* to within a few half units of a binade boundary, from
              * either side, in the grid of the value or of the binade
              * below */
    if (ad == 0) return ldexp(1.0, (int)(rnd() % 20) - 10);
    const double lo = ldexp(1.0, e - 1), hi = ldexp(1.0, e);
    const double target = (rnd() & 1) ? hi : lo;
    const double step = (rnd() & 1) ? ulp : ulp * 0.5;
    const double sg = d < 0 ? -1.0 : 1.0;
    return sg * (target - ad) + (double)((int)(rnd() % 9) - 4) * step * 0.5;
  }
  }
}

int main(int argc, char **argv)
{
  const long runs = argc > 1 ? atol(argv[1]) : 100000;
  rng_s = 0x9E3779B97F4A7C15ull ^ (uint64_t)(argc > 2 ? atol(argv[2]) : 1);
  long mism = 0, first_bad = -1;
  double t[64], out[64], seq[64];
  for (long run = 0; run < runs; run++) {
    /* a start value: zero, a subnormal, or a value of some binade */
    double d0;
    switch (rnd() % 5) {
    case 0: d0 = 0.0; break;
    case 1: d0 = ldexp((double)(rnd() % 100), -1074); break;
    case 2: d0 = -ldexp(0.5 + unif(), (int)(rnd() % 40) - 20); break;
    default: d0 = ldexp(0.5 + unif(), (int)(rnd() % 60) - 30); break;
    }
    const int m = 1 + (int)(rnd() % 64);
    const int mix = (int)(rnd() % 10);
    double dd = d0;
    for (int l = 0; l < m; l++) {
      const int kind = (rnd() % 3 == 0) ? (int)(rnd() % 10) : mix;
      t[l] = term(dd, kind);
      dd = dd + t[l];
    }
    double ds = d0;
    for (int l = 0; l < m; l++) {
      ds = ds + t[l];
      seq[l] = ds;
    }
    double df = d0;
    fold64_emu(&df, t, m, out);
    int bad = memcmp(&df, &ds, 8) != 0;
    for (int l = 0; l < m; l++) bad |= memcmp(&out[l], &seq[l], 8) != 0;
    if (bad) {
      mism++;
      if (first_bad < 0) first_bad = run;
    }
  }
  printf("{\"runs\": %ld, \"mismatches\": %ld, \"first_bad\": %ld, \"rounds\": %lld, \"plain_steps\": %lld}\n", runs,
         mism, first_bad, (long long)rounds_total, (long long)plain_total);
  return mism ? 1 : 0;
}
