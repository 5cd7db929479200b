/*
 * fold64.h -- test-only: the rules of an FP64 running-sum fold, replayed
 * on the host by fold64_check.c.  Not used by the kernels: folding a
 * wavefront's 64 terms of a running sum this way costs about five times a
 * dependent FP64 addition per term on gfx950 (the 64-bit lane prefix sum
 * dominates a round), so HaloProfiles' running sums stay one lane each.
 *
 * A running sum d_j = d_{j-1} + t_j in FP64, rounded to nearest even at
 * every step, is AHF's order of accumulation (HaloProfiles' running sums,
 * ahf_halos.c:3961-5060).  While d stays inside one binade [2^e, 2^(e+1))
 * (either sign), d = K * 2^(e-52) with K an integer in [2^52, 2^53), and a
 * step whose exact result stays inside the binade with a margin of one
 * unit rounds on the same grid: K becomes K + r with r = round(t / 2^(e-52))
 * to nearest, where only an exact half needs the parity of K + floor(x) to
 * pick the even result.  So 64 steps of a wavefront are one integer prefix
 * sum of the lanes' r, the ties settled in lane order; the first lane whose
 * step leaves the binade (or starts from zero, a subnormal, or a term that
 * is not finite) takes the hardware addition, and the fold resumes after
 * it.
 */
#ifndef AHF_FOLD64_H
#define AHF_FOLD64_H

#include <math.h>
#include <stdint.h>

#define AHF_HD static inline

/* d as sign, binade e and K in [2^52, 2^53); 0 for zero, subnormals, inf
 * and NaN */
AHF_HD int ahf_f64_binade(double d, int *e, int64_t *K, int *neg)
{
  union {
    double f;
    uint64_t u;
  } b;
  b.f = d;
  const int ex = (int)((b.u >> 52) & 0x7FFu);
  if (ex == 0 || ex == 0x7FF) return 0;
  *e = ex - 1023;
  *neg = (int)(b.u >> 63);
  *K = (int64_t)((b.u & ((1ull << 52) - 1)) | (1ull << 52));
  return 1;
}

/* one term t against the binade (e, neg): x = t / 2^(e-52), sign-folded so
 * that the sum grows K; q = floor(x), up: the fraction is above one half,
 * tie: it is exactly one half.  0 when t is not finite or |x| >= 2^51 (the
 * step leaves the binade). */
AHF_HD int ahf_f64_lane(double t, int e, int neg, int64_t *q, int *up, int *tie)
{
  *q = 0;
  *up = 0;
  *tie = 0;
  if (!(t == t) || t > 1e300 || t < -1e300) return 0;
  double x = ldexp(t, 52 - e);
  if (neg) x = -x;
  if (!(x < 2251799813685248.0 && x > -2251799813685248.0)) return 0;
  const double fq = floor(x);
  const double f = x - fq; /* exact */
  *q = (int64_t)fq;
  *up = f > 0.5;
  *tie = f == 0.5;
  return 1;
}

/* the exact sum Kb + x, x in [q, q + 1), lies in [2^52 + 1, 2^53 - 1]: its
 * rounding stays on the binade's grid */
AHF_HD int ahf_f64_in(int64_t Kb, int64_t q)
{
  const int64_t lo = Kb + q;
  return lo >= ((int64_t)1 << 52) + 1 && lo <= ((int64_t)1 << 53) - 2;
}

/* the value of K in binade (e, neg) */
AHF_HD double ahf_f64_make(int64_t K, int e, int neg)
{
  const double v = ldexp((double)K, e - 52);
  return neg ? -v : v;
}

#endif /* AHF_FOLD64_H */
