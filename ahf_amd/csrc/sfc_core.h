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
 * The key ranges the gather walks for a halo: slot s of the 3x3x3 block
 * (x offset outer, z inner, each -1, 0, +1 around the centre's cell,
 * periodic) holds [kmin, kmax] at full resolution, or kmin > kmax when the
 * cell is too far to hold a member (ahf_halos_sfc.c:254-296).  On the
 * 2-cell grid (bits = 1) the 8 cells are walked in key order 0..7 with no
 * distance test, and slots 8..26 are empty.  Returns bits.
 */
AHF_HD unsigned ahf_gather_cells(double cx, double cy, double cz, double gatherRad, uint64_t kmin[AHF_SHELL],
                                 uint64_t kmax[AHF_SHELL])
{
  const unsigned bits = ahf_gather_bits(gatherRad);
  const unsigned sh = 3u * (AHF_SFC_BITS - bits);
  const uint64_t span = ((uint64_t)1 << sh) - 1;
  for (int s = 0; s < AHF_SHELL; s++) {
    kmin[s] = 1;
    kmax[s] = 0;
  }
  if (bits == 1) {
    for (uint64_t k = 0; k < 8; k++) {
      kmin[k] = k << sh;
      kmax[k] = (k << sh) + span;
    }
    return bits;
  }
  const uint64_t n1 = (uint64_t)1 << bits;
  const uint64_t bx = ahf_grid_coord(cx, bits), by = ahf_grid_coord(cy, bits), bz = ahf_grid_coord(cz, bits);
  const double fac = 1. / (double)n1;
  const double reach = 0.5 * sqrt(3.) * fac + AHF_GATHERRAD_FAC * gatherRad;
  int s = 0;
  for (uint64_t i = n1 - 1; i <= n1 + 1; i++) {
    const uint64_t x = (bx + i) % n1;
    for (uint64_t j = n1 - 1; j <= n1 + 1; j++) {
      const uint64_t y = (by + j) % n1;
      for (uint64_t k = n1 - 1; k <= n1 + 1; k++, s++) {
        const uint64_t z = (bz + k) % n1;
        const double dx = ahf_gather_sep((fac * (double)x) + 0.5 * fac, cx);
        const double dy = ahf_gather_sep((fac * (double)y) + 0.5 * fac, cy);
        const double dz = ahf_gather_sep((fac * (double)z) + 0.5 * fac, cz);
        if (sqrt(dx * dx + dy * dy + dz * dz) > reach) continue;
        const uint64_t key = ahf_hilbert_key(x, y, z, bits);
        kmin[s] = key << sh;
        kmax[s] = (key << sh) + span;
      }
    }
  }
  return bits;
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
