/*
 * sfc_core.h -- AHF's Hilbert keys and the cell geometry of the halo gather,
 * shared by the device kernels (hipcc) and host C (dropin.c, snapgen.c).
 *
 * Keys: src/libsfc/hilbert_util.c:68-92 (calcHikey) over the curve of
 * src/libsfc/hilbert.c:197-241, at BITS_PER_DIMENSION = 21 bits per axis
 * (src/param.h:92); AHF orders its particles by them (src/main.c:341-356).
 *
 * Gather: ahf_halos_sfc_gatherParts (src/libahf/ahf_halos_sfc.c:171-367)
 * takes the particles within gatherRad of the centre from the cells of a
 * 3x3x3 block around the centre's cell, on the coarsest grid whose cells
 * are wider than twice 1.001*gatherRad; the cells in the block's order, the
 * particles of a cell in key order.  A cell is a contiguous key range of
 * the sorted particles.
 */
#ifndef AHF_SFC_CORE_H
#define AHF_SFC_CORE_H

#include <stdint.h>
#if !defined(__HIPCC__)
#include <math.h>
#endif

#ifndef AHF_HD
#if defined(__HIPCC__)
#define AHF_HD __host__ __device__ inline
#else
#define AHF_HD static inline
#endif
#endif

#define AHF_SFC_BITS 21          /* BITS_PER_DIMENSION, src/param.h:92 */
#define AHF_GATHERRAD_FAC 1.001  /* src/libahf/ahf_halos_sfc.c:46 */
#define AHF_SHELL 27             /* cells of the search block */

/*
 * Hilbert index of cell (x, y, z) of a 2^bits grid.  Level by level from
 * the top: the 3-bit digit (x bit 0, y bit 1, z bit 2) XOR the digit of the
 * level above, XOR the flip bit, rotated right by the current rotation; the
 * rotation then advances by one, plus one or two when the digit's lowest
 * set bit is bit 0 or bit 1.  Every digit but the top one then has bit 2
 * flipped, and the index is Gray-decoded (each bit XOR all bits above it).
 */
AHF_HD uint64_t ahf_hilbert_key(uint64_t x, uint64_t y, uint64_t z, unsigned bits)
{
  uint64_t h = 0;
  unsigned rot = 0, flip = 0, above = 0;
  for (int b = (int)bits - 1; b >= 0; b--) {
    const unsigned d = (unsigned)(((x >> b) & 1u) | (((y >> b) & 1u) << 1) | (((z >> b) & 1u) << 2));
    const unsigned v = (d ^ above ^ flip) & 7u;
    above = d;
    const unsigned t = ((v >> rot) | (v << (3u - rot))) & 7u;
    h = (h << 3) | t;
    flip = 1u << rot;
    rot += (t & 1u) ? 2u : (t & 2u) ? 3u : 1u;
    if (rot >= 3u) rot -= 3u;
  }
  uint64_t top4 = 0;
  for (unsigned l = 1; l < bits; l++) top4 |= (uint64_t)4 << (3 * (bits - 1 - l));
  h ^= top4;
  h ^= h >> 1;
  h ^= h >> 2;
  h ^= h >> 4;
  h ^= h >> 8;
  h ^= h >> 16;
  h ^= h >> 32;
  return h;
}

/* the grid cell of a position in box units on a 2^bits grid, clamped to
 * the last cell (hilbert_util.c:75-84) */
AHF_HD uint64_t ahf_grid_coord(double x, unsigned bits)
{
  const uint64_t max = (uint64_t)1 << bits;
  uint64_t c = (uint64_t)trunc(x * (double)max);
  return c == max ? max - 1 : c;
}

AHF_HD uint64_t ahf_sfc_key(double x, double y, double z, unsigned bits)
{
  return ahf_hilbert_key(ahf_grid_coord(x, bits), ahf_grid_coord(y, bits), ahf_grid_coord(z, bits), bits);
}

/* search grid of the gather: the finest level (at most 21) whose cells are
 * wider than twice 1.001*gatherRad (ahf_halos_sfc.c:240-252) */
AHF_HD unsigned ahf_gather_bits(double gatherRad)
{
  unsigned bits = 1;
  while ((AHF_GATHERRAD_FAC * gatherRad < 1. / (double)(1 << (bits + 1))) && (bits + 1) <= AHF_SFC_BITS) bits++;
  return bits;
}

/* periodic separation of one axis as the gather measures it
 * (LOCAL_CAL_DXYZ, ahf_halos_sfc.c:66-75) */
AHF_HD double ahf_gather_sep(double a, double b)
{
  double d = fabs(a - b);
  if (d > 0.5) d = 1.0 - d;
  return d;
}

/*
 * The key range the gather walks for slot s of a halo's 3x3x3 search block
 * (x offset outer, z inner, each -1, 0, +1 around the centre's cell,
 * periodic): [*kmin, *kmax] at full resolution, or *kmin > *kmax when the
 * cell is too far to hold a member (ahf_halos_sfc.c:254-296).  On the
 * 2-cell grid (bits = 1) slots 0..7 are the 8 cells in key order, with no
 * distance test, and slots 8..26 are empty.
 */
AHF_HD void ahf_gather_cell(double cx, double cy, double cz, double gatherRad, int s, uint64_t *kmin, uint64_t *kmax)
{
  const unsigned bits = ahf_gather_bits(gatherRad);
  const unsigned sh = 3u * (AHF_SFC_BITS - bits);
  const uint64_t span = ((uint64_t)1 << sh) - 1;
  *kmin = 1;
  *kmax = 0;
  if (bits == 1) {
    if (s < 8) {
      *kmin = (uint64_t)s << sh;
      *kmax = ((uint64_t)s << sh) + span;
    }
    return;
  }
  const uint64_t n1 = (uint64_t)1 << bits;
  const double fac = 1. / (double)n1;
  const double reach = 0.5 * sqrt(3.) * fac + AHF_GATHERRAD_FAC * gatherRad;
  const uint64_t x = (ahf_grid_coord(cx, bits) + n1 - 1 + (uint64_t)(s / 9)) % n1;
  const uint64_t y = (ahf_grid_coord(cy, bits) + n1 - 1 + (uint64_t)((s / 3) % 3)) % n1;
  const uint64_t z = (ahf_grid_coord(cz, bits) + n1 - 1 + (uint64_t)(s % 3)) % n1;
  const double dx = ahf_gather_sep((fac * (double)x) + 0.5 * fac, cx);
  const double dy = ahf_gather_sep((fac * (double)y) + 0.5 * fac, cy);
  const double dz = ahf_gather_sep((fac * (double)z) + 0.5 * fac, cz);
  if (sqrt(dx * dx + dy * dy + dz * dz) > reach) return;
  const uint64_t key = ahf_hilbert_key(x, y, z, bits);
  *kmin = key << sh;
  *kmax = (key << sh) + span;
}

/* the member test: squared periodic distance of an FP32 position from the
 * centre against gatherRad^2 (ahf_halos_sfc.c:327-331) */
AHF_HD int ahf_gather_member(float px, float py, float pz, double cx, double cy, double cz, double gatherRad2)
{
  const double dx = ahf_gather_sep((double)px, cx), dy = ahf_gather_sep((double)py, cy),
               dz = ahf_gather_sep((double)pz, cz);
  const double d2 = dx * dx + dy * dy + dz * dz;
  return d2 <= gatherRad2;
}

#endif
