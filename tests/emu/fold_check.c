/*
 * fold_check.c -- host replay of the wavefront fold of the heavy deposit
 * kernels (ahf_amd/csrc/amr_gpu.hip, fold_wave) against AHF's sequential
 * accumulation d = (float)((double)d + c), on random and adversarial runs:
 * starts at -1, at zero, on binade edges and subnormals; contributions that
 * are exact dyadic fractions (ties at every scale), zeros, values that cross
 * binades and zero, huge and tiny values, and negative ones.  The 64 lanes of
 * a wavefront are arrays here; the per-lane arithmetic is the same header
 * code the kernels use.  Exit status 0 when every result is bit-identical.
 *
 * usage: fold_check [runs] [seed]
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../ahf_amd/csrc/amr_core.h"

static float fold_seq(float d, const double *c, int n)
{
  for (int i = 0; i < n; i++) d = (float)((double)d + c[i]);
  return d;
}

/* the kernel's fold of lanes [0, mm), replayed lane by lane */
static float fold_wave(float d, const double *cv, int mm, long *stats)
{
  int start = 0;
  while (start < mm) {
    int ok;
    const int e = ahf_fold_exponent(d, &ok);
    if (!ok) {
      d = (float)((double)d + cv[start]);
      start++;
      stats[0]++;
      continue;
    }
    const int neg = d < 0.0f;
    const int64_t K0 = ahf_fold_K(d, e);
    int64_t R[64], r[64];
    int tie[64], lok[64];
    for (int l = start; l < mm; l++) lok[l] = ahf_fold_lane(cv[l], e, &R[l], &r[l], &tie[l]);
    for (int l = start; l < mm; l++)
      if (lok[l] && tie[l]) {
        int64_t k = K0;
        for (int j = start; j < l; j++) k += r[j];
        if ((k + r[l]) & 1) r[l] += 1;
      }
    /* steps clamped to +-2^24 and summed in 32 bits, as the kernel does */
    int32_t P = 0;
    int bad = mm;
    for (int l = start; l < mm; l++) {
      const int64_t Kin = K0 + (int64_t)P;
      if (!(lok[l] && ahf_fold_K_in(Kin, neg) && ahf_fold_T_in(Kin, R[l], neg))) {
        bad = l;
        break;
      }
      const int64_t lim = (int64_t)1 << 24;
      P += (int32_t)(r[l] > lim ? lim : (r[l] < -lim ? -lim : r[l]));
    }
    if (bad > start) d = (float)ldexp((double)(K0 + P), e - 23);
    stats[1]++;
    if (bad < mm) {
      d = (float)((double)d + cv[bad]);
      stats[0]++;
      start = bad + 1;
    } else {
      start = mm;
    }
  }
  return d;
}

static uint64_t rs = 88172645463325252ull;
static uint64_t xr(void)
{
  rs ^= rs << 13;
  rs ^= rs >> 7;
  rs ^= rs << 17;
  return rs;
}
static double ur(void) { return (double)(xr() >> 11) * (1.0 / 9007199254740992.0); }

static double draw_c(int kind)
{
  switch (kind) {
  case 0: /* TSC-like: products of weights in [0, 0.75] times a mass scale */
    return (0.75 * ur()) * (0.75 * ur()) * (0.75 * ur()) * ldexp(1.0, -(int)(xr() % 20));
  case 1: /* exact dyadic fractions: ties at many scales */
    return ldexp((double)(xr() % 64), -(int)(xr() % 40));
  case 2: /* wide dynamic range */
    return ldexp(ur(), (int)(xr() % 80) - 60);
  case 3: /* zeros and tiny values */
    return (xr() & 1) ? 0.0 : ldexp(ur(), -1000 - (int)(xr() % 60));
  case 4: /* negative and positive */
    return (ur() - 0.5) * ldexp(1.0, -(int)(xr() % 30));
  default: /* half-ulp multiples of the running value's scale */
    return ldexp(0.5 + (double)(xr() % 3), -24 - (int)(xr() % 6));
  }
}

static float draw_start(int kind)
{
  switch (kind) {
  case 0: return -1.0f;
  case 1: return 0.0f;
  case 2: return ldexpf(1.0f, (int)(xr() % 40) - 20);
  case 3: return -ldexpf(1.0f, (int)(xr() % 40) - 20);
  case 4: return ldexpf(1.0f, -130);  /* subnormal */
  default: return (float)((ur() - 0.5) * 4.0);
  }
}

int main(int argc, char **argv)
{
  const long runs = argc > 1 ? atol(argv[1]) : 200000;
  if (argc > 2) rs ^= (uint64_t)atoll(argv[2]) * 0x9E3779B97F4A7C15ull;
  long stats[2] = {0, 0}, bad = 0;
  double c[64];
  for (long t = 0; t < runs; t++) {
    const int mm = 1 + (int)(xr() % 64);
    const int kind = (int)(xr() % 6);
    const int mixed = (xr() % 4) == 0;
    for (int l = 0; l < mm; l++) c[l] = draw_c(mixed ? (int)(xr() % 6) : kind);
    if (kind == 4 && !mixed)
      for (int l = 0; l < mm; l++) c[l] = fabs(c[l]) * (xr() % 8 ? 1.0 : -1.0);
    float d = draw_start((int)(xr() % 6));
    /* chain several chunks, as a heavy node does */
    const int chunks = 1 + (int)(xr() % 4);
    float a = d, b = d;
    for (int k = 0; k < chunks; k++) {
      a = fold_seq(a, c, mm);
      b = fold_wave(b, c, mm, stats);
      uint32_t ua, ub;
      memcpy(&ua, &a, 4);
      memcpy(&ub, &b, 4);
      if (ua != ub) {
        if (bad < 10)
          fprintf(stderr, "mismatch run %ld chunk %d kind %d mm %d: seq %.9g (%08x) wave %.9g (%08x)\n", t, k, kind,
                  mm, (double)a, ua, (double)b, ub);
        bad++;
        break;
      }
      for (int l = 0; l < mm; l++) c[l] = draw_c(mixed ? (int)(xr() % 6) : kind);
    }
  }
  printf("{\"runs\": %ld, \"mismatches\": %ld, \"scalar_steps\": %ld, \"parallel_rounds\": %ld}\n", runs, bad, stats[0],
         stats[1]);
  return bad ? 1 : 0;
}
