/*
 * march_rows_check.c -- host check of the march kernel's rows walk
 * (ahf_amd/csrc/amr_gpu.hip, k_gather_march, the rows fast path): for
 * random entry counts of a staged 18x18 source plane and lanes inside the
 * box and at the periodic faces, the entries the walk visits with the
 * functions it uses (ahf_row_start / ahf_row_next / ahf_row_src) must be
 * the sources in AHF's traversal order (ahf_axis_order on y and x), each
 * source's entries ascending, with the source's x offset.  Prints the
 * number of lanes whose walk differs; 0 passes.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../ahf_amd/csrc/amr_core.h"

#define W 18
#define N (W * W)

int main(void)
{
  uint32_t cnt[N], off[N + 1];
  uint64_t want[64], got[64];
  int bad = 0;
  srand(7);
  for (int trial = 0; trial < 200000; trial++) {
    const int dense = trial % 3;
    for (int u = 0; u < N; u++) cnt[u] = (uint32_t)(rand() % (dense == 0 ? 2 : (dense == 1 ? 3 : 6)));
    off[0] = 0;
    for (int u = 0; u < N; u++) off[u + 1] = off[u] + cnt[u];
    const int64_t L = 32;
    const int lx = rand() % 16, ly = rand() % 16;
    const int64_t X = (rand() % 3 == 0) ? 0 : (rand() % 2 ? L - 1 : 5);
    const int64_t Y = (rand() % 3 == 0) ? 0 : (rand() % 2 ? L - 1 : 7);
    int ox[3], oy[3];
    ahf_axis_order(X, L, ox);
    ahf_axis_order(Y, L, oy);
    int na = 0, nb = 0;
    for (int t = 0; t < 9; t++) {
      const int ey = oy[t / 3], ex = ox[t % 3];
      const int u = (ly + 1 + ey) * W + (lx + 1 + ex);
      for (uint32_t j = 0; j < cnt[u] && na < 64; j++)
        want[na++] = (uint64_t)(off[u] + j) << 8 | (uint64_t)(ey + 1) << 4 | (uint64_t)(ex + 1);
    }
    const int rot = ox[0] + 1;
    uint32_t ntot = 0;
    for (int bb = 0; bb < 3; bb++) {
      const int u = (ly + 1 + oy[bb]) * W + lx;
      ntot += off[u + 3] - off[u];
    }
    uint32_t q = 0, rem = 0, lo = 0, hi = 0, b1 = 0, b2 = 0;
    int br = -1, ey = 0;
    for (uint32_t it = 0; it < ntot; it++) {
      while (rem == 0) {
        br++;
        ey = oy[br];
        const int u = (ly + 1 + ey) * W + lx;
        lo = off[u];
        b1 = off[u + 1];
        b2 = off[u + 2];
        hi = off[u + 3];
        rem = hi - lo;
        q = ahf_row_start(lo, b1, b2, hi, rot);
      }
      if (nb < 64) got[nb++] = (uint64_t)q << 8 | (uint64_t)(ey + 1) << 4 | (uint64_t)ahf_row_src(q, b1, b2);
      q = ahf_row_next(q, lo, hi);
      rem--;
    }
    int same = na == nb;
    for (int i = 0; same && i < na; i++) same = want[i] == got[i];
    bad += !same;
  }
  printf("%d\n", bad);
  return bad != 0;
}
