/*
 * nr_order.h -- the per-segment rules of the level-synchronous replay of
 * Numerical Recipes' indexx (src/libutility/general.c:1093-1153), which AHF
 * uses to sort a halo's members by distance (src/libahf/ahf_halos.c:5846).
 *
 * indexx is a quicksort with an explicit stack: median of three at l, l+1,
 * ir (after moving the middle element to l+1), pivot a = arr[indx[l]], a
 * partition by two scans that swap the pairs they meet, the pivot put in
 * place, and straight insertion for segments of fewer than 8 elements.  Its
 * segments are disjoint, so the order in which the stack visits them does
 * not change the result, and every segment of a level can be partitioned
 * at once.  With, in the array as the median of three left it,
 *   P = the positions p in [l+2, ir] with arr >= a, ascending,
 *   Q = the positions q in [l, ir-1] with arr <= a, descending,
 * the scans meet P[k] and Q[k] at step k while P[k] < Q[k] and swap them
 * (P[k] - Q[k] grows with k, so these are the first K steps).  A step with
 * P[K] == Q[K] swaps an element with itself and goes on.  The step that
 * ends the loop scans past no swapped position: i stops at the smallest
 * swapped Q position, j at the largest swapped P position, so
 *   i = min(P[K'], Q[K-1]),  j = max(Q[K'], P[K-1])
 * with K' = K + 1 after a self swap and K' = K otherwise.  The element at
 * j then takes position l and the pivot position j; [l, j-1] and [i, ir]
 * are the next segments.  P holds ir and Q holds l (the median of three
 * puts arr[ir] >= a at ir and the pivot at l), so every list entry used
 * above exists.
 *
 * Positions are global over a flat array of several haloes; a segment
 * carries the global position B of its halo's first element, indexx's
 * 1-based indices are local to the halo (idx values 1..n) and key[] holds
 * the halo's keys in list order at B.. .  Shared by the HIP kernels
 * (unbind_gpu.hip) and the host replay of tests/emu/nr_check.c.
 */
#ifndef AHF_NR_ORDER_H
#define AHF_NR_ORDER_H

#include <stdint.h>

#if defined(__HIPCC__)
#define NR_HD __host__ __device__ inline
#else
#define NR_HD static inline
#endif

#define NR_SMALL 7 /* indexx's M: straight insertion when ir - l < M */

/* arr[indx[pos]] of the segment's halo */
NR_HD double nr_key(const double *key, const uint32_t *idx, int64_t B, int64_t pos)
{
  return key[B + (int64_t)idx[pos] - 1];
}

/* straight insertion of [L, R] (indexx's branch for ir - l < M).  indexx
 * walks down to index 1, but everything left of a segment is <= every
 * element in it (the partitions' invariant), so its walk stops at L - 1
 * at the latest; the walk here stops at L. */
NR_HD void nr_insertion(const double *key, uint32_t *idx, int64_t B, int64_t L, int64_t R)
{
  for (int64_t j = L + 1; j <= R; j++) {
    const uint32_t it = idx[j];
    const double a = key[B + (int64_t)it - 1];
    int64_t i = j - 1;
    for (; i >= L; i--) {
      if (nr_key(key, idx, B, i) <= a) break;
      idx[i + 1] = idx[i];
    }
    idx[i + 1] = it;
  }
}

/* the median of three of [L, R] (R - L >= M): the middle element (in the
 * halo's 1-based numbering) to L+1, then three compare-swaps.  Returns the
 * pivot key; *pv receives the pivot's index. */
NR_HD double nr_median3(const double *key, uint32_t *idx, int64_t B, int64_t L, int64_t R, uint32_t *pv)
{
  const int64_t lo = L - B + 1, hi = R - B + 1;
  const int64_t k = B + ((lo + hi) >> 1) - 1;
  uint32_t t;
  t = idx[k]; idx[k] = idx[L + 1]; idx[L + 1] = t;
  if (nr_key(key, idx, B, L + 1) > nr_key(key, idx, B, R)) { t = idx[L + 1]; idx[L + 1] = idx[R]; idx[R] = t; }
  if (nr_key(key, idx, B, L) > nr_key(key, idx, B, R)) { t = idx[L]; idx[L] = idx[R]; idx[R] = t; }
  if (nr_key(key, idx, B, L + 1) > nr_key(key, idx, B, L)) { t = idx[L + 1]; idx[L + 1] = idx[L]; idx[L] = t; }
  *pv = idx[L];
  return nr_key(key, idx, B, L);
}

/* membership of position g (key v) of segment [L, R] with pivot key a in P
 * (bit 0) and Q (bit 1) */
NR_HD unsigned nr_flags(double v, double a, int64_t g, int64_t L, int64_t R)
{
  return (unsigned)(g >= L + 2 && v >= a) | ((unsigned)(g <= R - 1 && v <= a) << 1);
}

/* where the scans stop, from P[0..nP) and Q[0..nQ) (global positions):
 * *K the swaps, *i and *j the final scan positions, *j_swapped 1 when j is
 * the swapped position P[K-1] (it then holds the element from Q[K-1]).
 * Returns 0, or -1 if a list runs short (impossible for a segment laid out
 * by nr_median3). */
NR_HD int nr_meet(const uint32_t *P, const uint32_t *Q, int64_t nP, int64_t nQ, int64_t *K, int64_t *i, int64_t *j,
                  int *j_swapped)
{
  int64_t lo = 0, hi = nP < nQ ? nP : nQ;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (P[mid] < Q[mid]) lo = mid + 1; else hi = mid;
  }
  const int64_t k = lo;
  const int self = k < nP && k < nQ && P[k] == Q[k];
  const int64_t kp = k + (self ? 1 : 0);
  if (kp >= nP || kp >= nQ) return -1;
  int64_t ii = P[kp], jj = Q[kp];
  int js = 0;
  if (k >= 1) {
    if ((int64_t)Q[k - 1] < ii) ii = Q[k - 1];
    /* P[k-1] may also be Q[K'] (an element equal to the pivot): it is the
     * swapped position either way */
    if ((int64_t)P[k - 1] >= jj) {
      jj = P[k - 1];
      js = 1;
    }
  }
  *K = k;
  *i = ii;
  *j = jj;
  *j_swapped = js;
  return 0;
}

#endif
